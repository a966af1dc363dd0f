// sgipc_linear_system.hpp -- header-only C++ facade over the sgipc C ABI with
// the shape of Stiff-GIPC's linear-system classes (namespace gipc,
// StiffGIPC/linear_system/): one GlobalLinearSystem that owns the solver
// configuration and the preconditioner, a triplet matrix the caller
// assembles into, and solve_linear_system() returning the PCG iteration
// count.  Errors throw sgipc::Error (the C ABI returns status codes; the
// reference aborts).
//
// Mapping to the reference:
//   GlobalLinearSystem            global_linear_system.h:24-66
//   solve_linear_system()         global_linear_system.cu:124-133
//   PCGSolverConfig               solver/pcg_solver.h:6-16
//   TripletMatrix                 GIPCTripletMatrix, linear_system/global_matrix.h:18-203
//   set_mas_partition()           setMAS_partition + initPreconditioner_Neighbor,
//                                 gl_main.cu:1066-1093, MASPreconditioner.cu:2264-2302
//   kinetic/fem/barrier/mass/rhs  GIPC::computeGradientAndHessian, GIPC.cu:10110-10344
#ifndef SGIPC_LINEAR_SYSTEM_HPP
#define SGIPC_LINEAR_SYSTEM_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sgipc.h"

namespace sgipc {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const sgipc_ctx* ctx, const char* where) {
    if (rc == SGIPC_OK) return;
    std::string msg = where;
    msg += ": ";
    msg += ctx ? sgipc_last_error(ctx) : "context creation failed";
    throw Error(rc, msg);
}

// PCGSolverConfig (solver/pcg_solver.h:6-16) plus this build's options
struct PCGSolverConfig {
    double global_tol_rate = 1e-4;           // pcg_threshold (parameterSetting.txt:14)
    double max_iter_ratio = 0.3;             // max iterations = ratio * dof
    int max_iter = 0;                        // > 0 overrides the ratio
    int stop_rule = SGIPC_STOP_REF_RZ;       // reference |rz| <= tol rz0 test
    int precision = SGIPC_PRECISION_FP64;
    int check_every = 8;
};

// View of the caller's device triplet arrays (GIPCTripletMatrix).  Assembly
// calls append at global_triplet_offset and advance it, as GIPC does.
struct TripletMatrix {
    double* values = nullptr;  // 9 doubles per block, column-major
    int* rows = nullptr;
    int* cols = nullptr;
    long long capacity = 0;
    long long global_triplet_offset = 0;
};

class GlobalLinearSystem {
public:
    explicit GlobalLinearSystem(int device = 0, void* hip_stream = nullptr) {
        check(sgipc_ctx_create(device, &ctx_), nullptr, "sgipc_ctx_create");
        if (hip_stream) check(sgipc_ctx_set_stream(ctx_, hip_stream), ctx_, "sgipc_ctx_set_stream");
    }
    ~GlobalLinearSystem() {
        if (ctx_) sgipc_ctx_destroy(ctx_);
    }
    GlobalLinearSystem(const GlobalLinearSystem&) = delete;
    GlobalLinearSystem& operator=(const GlobalLinearSystem&) = delete;

    sgipc_ctx* handle() const { return ctx_; }
    PCGSolverConfig& config() { return cfg_; }
    const sgipc_solve_stats& last_stats() const { return stats_; }
    TripletMatrix& triplets() { return trip_; }

    // ---- one-time setup --------------------------------------------------
    void set_fem_mesh(int n, int T, const uint32_t* h_tets) {
        check(sgipc_set_fem_mesh(ctx_, n, T, h_tets), ctx_, "sgipc_set_fem_mesh");
        n_fem_ = n;
    }
    // Vertex neighbours from the tets and bank slot maps from ascending part
    // ids (host), then the MAS topology upload.
    void set_mas_partition(int n, int T, const uint32_t* h_tets, const int* h_part_id, int bank = 16) {
        std::vector<uint32_t> start(n > 0 ? n : 1), num(n > 0 ? n : 1), list(12 * (size_t)T + 1);
        int len = 0;
        check(sgipc_build_vertex_neighbors(n, T, h_tets, start.data(), num.data(), list.data(), (int)list.size(),
                                           &len),
              ctx_, "sgipc_build_vertex_neighbors");
        const int cap = n > 0 ? bank * (h_part_id[n - 1] + 1) : 0;
        std::vector<int> part_map(cap > 0 ? cap : 1), real_map(n > 0 ? n : 1);
        int M = 0;
        check(sgipc_build_bank_map(n, h_part_id, bank, part_map.data(), cap, &M, real_map.data()), ctx_,
              "sgipc_build_bank_map");
        // contact pairs carry vertex ids: FEM vertices follow the body vertices
        check(sgipc_set_mas_topology(ctx_, n, n_body_vertices_, start.data(), num.data(), list.data(), len,
                                     part_map.data(), M, real_map.data(), bank),
              ctx_, "sgipc_set_mas_topology");
    }
    // Affine bodies (ABDSystem / ABDSimData, abd_system/abd_sim_data.h:40-100):
    // call before set_mas_partition.  Body vertices come first in vertex
    // space; FEM block rows then start at 4 * n_bodies (abd_fem_count_info).
    void set_affine_bodies(int n_bodies, int n_body_vertices, const int* h_vstart, const double* h_xbar,
                           const double* h_mass12, const double* h_volume, const int* h_fixed = nullptr) {
        check(sgipc_abd_setup(ctx_, n_bodies, n_body_vertices, h_vstart, h_xbar, h_mass12, h_volume, h_fixed), ctx_,
              "sgipc_abd_setup");
        n_bodies_ = n_bodies;
        n_body_vertices_ = n_body_vertices;
    }
    void set_triplet_matrix(double* d_vals, int* d_rows, int* d_cols, long long capacity) {
        check(sgipc_triplets_view(ctx_, d_vals, d_rows, d_cols, capacity), ctx_, "sgipc_triplets_view");
        trip_ = TripletMatrix{d_vals, d_rows, d_cols, capacity, 0};
    }
    void set_preconditioner(int kind) { precond_ = kind; }
    // contact pairs of the current iterate (MAS contact connectivity)
    void set_contact_pairs(const int* d_pairs, int n_pairs) {
        d_pairs_ = d_pairs;
        n_pairs_ = n_pairs;
    }

    // ---- assembly: GIPC::computeGradientAndHessian ------------------------
    void begin_assembly() { trip_.global_triplet_offset = 0; }
    void kinetic_gradient(int n, const double* d_x, const double* d_xtilde, const double* d_mass, double* d_shape) {
        check(sgipc_kinetic_gradient(ctx_, n, d_x, d_xtilde, d_mass, d_shape), ctx_, "sgipc_kinetic_gradient");
    }
    void assemble_barrier(int n_pairs, const int* d_pairs, const int* d_mat_index, const int counts[3],
                          const double* d_x, const double* d_xrest, double dhat, double kappa, const int* d_btype,
                          double* d_fb) {
        check(sgipc_assemble_barrier(ctx_, n_pairs, d_pairs, d_mat_index, counts, d_x, d_xrest, dhat, kappa,
                                     d_btype, 0, trip_.global_triplet_offset, d_fb),
              ctx_, "sgipc_assemble_barrier");
        trip_.global_triplet_offset += 16LL * counts[2] + 9LL * counts[1] + 4LL * counts[0];
    }
    void assemble_fem(int T, const double* d_x, const uint32_t* d_tets, const double* d_DmInv, const double* d_vol,
                      const double* d_mu, const double* d_lambda, double dt, const int* d_btype, double* d_shape) {
        check(sgipc_assemble_fem(ctx_, T, d_x, d_tets, d_DmInv, d_vol, d_mu, d_lambda, dt, d_btype, 4 * n_bodies_,
                                 trip_.global_triplet_offset, d_shape),
              ctx_, "sgipc_assemble_fem");
        trip_.global_triplet_offset += 10LL * T;
    }
    void mass_diagonal(int n, const double* d_mass) {
        check(sgipc_mass_triplets(ctx_, n, d_mass, 4 * n_bodies_, trip_.global_triplet_offset), ctx_,
              "sgipc_mass_triplets");
        trip_.global_triplet_offset += n;
    }
    // collision detection (GIPC::buildCP, mlbvh.cu): LBVH candidates, then
    // the narrow phase into reference-encoded pairs and their class counts;
    // the caller sizes the buffers (SGIPC_ERR_CAPACITY reports the need)
    int detect_contacts(int n_surf, const uint32_t* d_surf, int n_faces, const uint32_t* d_faces, int n_edges,
                        const uint32_t* d_edges, const double* d_x, const double* d_xrest, const int* d_btype,
                        double dhat, int* d_pt, int* d_ee, int cand_capacity, int* d_pairs, int* d_mat_index,
                        int pair_capacity, int counts[3]) {
        int npt = 0, nee = 0, np = 0;
        check(sgipc_broad_phase(ctx_, n_surf, d_surf, n_faces, d_faces, n_edges, d_edges, d_x, d_btype, nullptr, dhat,
                                d_pt, cand_capacity, &npt, d_ee, cand_capacity, &nee),
              ctx_, "sgipc_broad_phase");
        check(sgipc_classify_contacts(ctx_, npt, d_pt, nee, d_ee, d_x, d_xrest, dhat, d_pairs, d_mat_index,
                                      pair_capacity, &np, counts),
              ctx_, "sgipc_classify_contacts");
        return np;
    }
    // line search: largest feasible step along -d_dir (GIPC::buildFullCP +
    // self_largestFeasibleStepSize, GIPC.cu:9085, 8880-8915)
    double ccd_step_bound(int n_surf, const uint32_t* d_surf, int n_faces, const uint32_t* d_faces, int n_edges,
                          const uint32_t* d_edges, const double* d_x, const double* d_dir, double dhat,
                          double slackness, int* d_pt, int* d_ee, int cand_capacity) {
        int npt = 0, nee = 0;
        double a = 1.0;
        check(sgipc_broad_phase_ccd(ctx_, n_surf, d_surf, n_faces, d_faces, n_edges, d_edges, d_x, d_dir, 1.0, nullptr,
                                    nullptr, dhat, d_pt, cand_capacity, &npt, d_ee, cand_capacity, &nee),
              ctx_, "sgipc_broad_phase_ccd");
        check(sgipc_ccd_step_bound(ctx_, npt, d_pt, nee, d_ee, d_x, d_dir, slackness, &a), ctx_,
              "sgipc_ccd_step_bound");
        return a;
    }
    // ground plane n.x = offset (GIPC::_groundNormal/_groundOffset):
    // close vertices of d_cand into d_close (candidate order), then their
    // barrier gradient and 3x3 blocks (GIPC.cu:6433-6514)
    int ground_detect(int n_cand, const uint32_t* d_cand, const double* d_x, const double normal[3], double offset,
                      double dhat, uint32_t* d_close) {
        int k = 0;
        check(sgipc_ground_detect(ctx_, n_cand, d_cand, d_x, normal, offset, dhat, d_close, &k), ctx_,
              "sgipc_ground_detect");
        return k;
    }
    void assemble_ground(int n_close, const uint32_t* d_close, const double* d_x, const double normal[3],
                         double offset, double dhat, double kappa, const int* d_btype, double* d_fb) {
        check(sgipc_assemble_ground(ctx_, n_close, d_close, d_x, normal, offset, dhat, kappa, d_btype, 0,
                                    trip_.global_triplet_offset, d_fb),
              ctx_, "sgipc_assemble_ground");
        trip_.global_triplet_offset += n_close;
    }
    // FEMLinearSubsystem::assemble: b = fb + shape, zero on pinned vertices
    // (with bodies: the FEM part, fb and b from the FEM vertex / row onwards)
    void assemble_rhs(int n, const double* d_fb, const double* d_shape, const int* d_btype, double* d_b) {
        check(sgipc_fem_rhs(ctx_, n, d_fb, d_shape, d_btype, d_b), ctx_, "sgipc_fem_rhs");
    }
    // ---- affine bodies: ABDSystem::setup_abd_system_gradient_hessian -----
    // the contact barrier over all vertices, written past the converted range
    // (end of the store) and mapped into DOF space at the current offset
    void assemble_barrier_with_bodies(int n_pairs, const int* d_pairs, const int* d_mat_index, const int counts[3],
                                      const double* d_x, const double* d_xrest, double dhat, double kappa,
                                      const int* d_btype, double* d_fb) {
        const long long m = 16LL * counts[2] + 9LL * counts[1] + 4LL * counts[0];
        const long long src = trip_.capacity - m;
        check(sgipc_assemble_barrier(ctx_, n_pairs, d_pairs, d_mat_index, counts, d_x, d_xrest, dhat, kappa, d_btype,
                                     0, src, d_fb),
              ctx_, "sgipc_assemble_barrier");
        long long n = 0;
        check(sgipc_abd_map_contacts(ctx_, src, m, trip_.global_triplet_offset, &n), ctx_, "sgipc_abd_map_contacts");
        trip_.global_triplet_offset += n;
    }
    // body gradient (the first 12 * n_bodies entries of b) and body blocks;
    // then the contact forces of the body vertices, J^T fb
    void assemble_bodies(const double* d_q, const double* d_q_tilde, double kappa_abd, double dt, const double* d_fb,
                         double* d_b) {
        check(sgipc_abd_assemble(ctx_, d_q, d_q_tilde, kappa_abd, dt, trip_.global_triplet_offset, d_b), ctx_,
              "sgipc_abd_assemble");
        trip_.global_triplet_offset += 10LL * n_bodies_;
        if (d_fb) check(sgipc_abd_contact_gradient(ctx_, d_fb, d_b), ctx_, "sgipc_abd_contact_gradient");
    }
    // body vertex positions from q, or their step from dq (cal_x_from_q.cu)
    void body_positions(const double* d_q, double* d_x) {
        check(sgipc_abd_positions(ctx_, d_q, d_x), ctx_, "sgipc_abd_positions");
    }
    double body_energy(const double* d_q, const double* d_q_tilde, double kappa_abd, double dt) {
        double e = 0.0;
        check(sgipc_abd_energy(ctx_, d_q, d_q_tilde, kappa_abd, dt, &e), ctx_, "sgipc_abd_energy");
        return e;
    }

    // ---- solve: GlobalLinearSystem::solve_linear_system -----------------
    // convert the assembled triplets in place, build the preconditioner, run
    // PCG; returns the iteration count (0 for an empty system).
    std::size_t solve_linear_system(const double* d_b, double* d_x) {
        const int rows = 4 * n_bodies_ + n_fem_;
        if (rows <= 0 || trip_.global_triplet_offset == 0) return 0;
        int K = 0;
        check(sgipc_convert(ctx_, trip_.global_triplet_offset, rows, &K), ctx_, "sgipc_convert");
        // MAS on the FEM rows; with bodies, their 12x12 block-Jacobi below
        check(sgipc_precond_build(ctx_, precond_, 4 * n_bodies_, n_fem_, d_pairs_, n_pairs_), ctx_,
              "sgipc_precond_build");
        sgipc_solve_opts o{cfg_.precision, cfg_.stop_rule, cfg_.global_tol_rate, cfg_.max_iter_ratio, cfg_.max_iter,
                           cfg_.check_every};
        check(sgipc_solve(ctx_, d_b, d_x, &o, &stats_), ctx_, "sgipc_solve");
        return (std::size_t)stats_.iterations;
    }
    // Independent scenes assembled at disjoint vertex ranges into the one
    // triplet store: scene s owns block rows [row_start[s], row_start[s+1]).
    // Every scene runs its own PCG in shared launches; per-scene stats in
    // `stats` (resized to the scene count).  No reference counterpart.
    void solve_linear_system_batched(const std::vector<int>& row_start, const double* d_b, double* d_x,
                                     std::vector<sgipc_solve_stats>& stats) {
        const int S = (int)row_start.size() - 1;
        if (S <= 0) throw Error(SGIPC_ERR_INVALID_ARG, "solve_linear_system_batched: no scenes");
        int K = 0;
        check(sgipc_convert(ctx_, trip_.global_triplet_offset, n_fem_, &K), ctx_, "sgipc_convert");
        check(sgipc_precond_build(ctx_, precond_, 0, n_fem_, d_pairs_, n_pairs_), ctx_, "sgipc_precond_build");
        sgipc_solve_opts o{cfg_.precision, cfg_.stop_rule, cfg_.global_tol_rate, cfg_.max_iter_ratio, cfg_.max_iter,
                           cfg_.check_every};
        stats.assign(S, sgipc_solve_stats{});
        check(sgipc_solve_batched(ctx_, S, row_start.data(), d_b, d_x, &o, stats.data()), ctx_,
              "sgipc_solve_batched");
    }
    // time the candidate schedules on the converted system, keep the fastest
    void autotune(const double* d_b, const std::vector<int>& row_start = {}) {
        const int S = row_start.empty() ? 0 : (int)row_start.size() - 1;
        check(sgipc_autotune(ctx_, d_b, S, S ? row_start.data() : nullptr, cfg_.precision, nullptr, nullptr), ctx_,
              "sgipc_autotune");
    }
    // helpers of the solver contract (i_linear_system_solver.h:43-53)
    void spmv(double a, const double* d_x, double b, double* d_y) {
        check(sgipc_spmv(ctx_, a, d_x, b, d_y), ctx_, "sgipc_spmv");
    }
    void apply_preconditioner(const double* d_r, double* d_z) {
        check(sgipc_precond_apply(ctx_, d_r, d_z), ctx_, "sgipc_precond_apply");
    }
    void synchronize() { check(sgipc_ctx_synchronize(ctx_), ctx_, "sgipc_ctx_synchronize"); }

private:
    sgipc_ctx* ctx_ = nullptr;
    PCGSolverConfig cfg_;
    sgipc_solve_stats stats_{};
    TripletMatrix trip_;
    int n_fem_ = 0;
    int n_bodies_ = 0, n_body_vertices_ = 0;
    int precond_ = SGIPC_PRECOND_MAS;
    const int* d_pairs_ = nullptr;
    int n_pairs_ = 0;
};

}  // namespace sgipc

#endif
