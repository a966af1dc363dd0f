// hybrid_step -- one Newton step of a hybrid scene (affine bodies over a FEM
// body), written the way Stiff-GIPC drives its linear system, against the
// reference-shaped C++ API (include/gipc_linear_system.hpp).
//
// The simulator side below (GIPC, device_TetraData, the ABD state, and the
// two subsystems that glue them to the linear system) is this example's own
// minimal stand-in for the reference's classes; GIPC::create_LinearSystem
// makes the same GlobalLinearSystem::create<T>() calls, with the same
// arguments, as StiffGIPC/gipc/gipc.cu:42-70, and the Newton step assembles
// into gipc_global_triplet and calls solve_linear_system() as
// GIPC::computeGradientAndHessian / calculateMovingDirection do
// (StiffGIPC/GIPC.cu:10110-10344, 10581-10592): contact barrier in vertex
// space mapped to DOF space, FEM kinetic and SNK1 terms, mass diagonal,
// body terms with the contact forces of the body vertices; the subsystems
// then supply the right-hand side and take the solution.
//
// usage: hybrid_step <scene.bin> <out.bin>
// scene.bin (little endian): "SGIPCHY1", int32 nb, nav, nf, T, bank, npairs,
// counts[3], stop_rule (unused: the reference stop rule applies), max_iter
// (unused), double dt, dhat, kappa, kappa_abd, tol (the PCG threshold),
// then vstart[nb+1] (int32), xbar[3 nav], mass12[144 nb], volume[nb]
// (double), fixed[nb] (int32), q[12 nb], q_tilde[12 nb], x[3 (nav+nf)],
// x_rest[3 (nav+nf)] (double), btype[nav+nf] (int32), x_tilde_fem[3 nf],
// mass_fem[nf] (double), tets[4T] (uint32, FEM-local ids), DmInv[9T],
// vol[T], mu[T], lam[T] (double), part[nf], pairs[4 npairs],
// mat_index[npairs] (int32).
// out.bin: int32 iterations, converged, double rel_residual, dx[12 nb + 3 nf].
// Exit codes: 0 ok, 2 bad input, 3 library error (message on stderr).
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <vector>

#include "gipc_linear_system.hpp"

namespace {

struct Reader {
    FILE* f;
    template <typename T>
    void get(std::vector<T>& v, size_t n) {
        v.resize(n);
        if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("truncated scene file");
    }
    template <typename T>
    T one() {
        T v;
        if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("truncated scene file");
        return v;
    }
};

struct Scene {
    int nb, nav, nf, T, bank, npairs, counts[3];
    double dt, dhat, kappa, kappa_abd, tol;
    std::vector<int> vstart, fixed, btype, part, pairs, mat_index;
    std::vector<double> xbar, mass12, volume, q, qt, x, xrest, xt, mass, DmInv, vol, mu, lam;
    std::vector<uint32_t> tets;
};

struct DeviceArena {
    sgipc_ctx* ctx;
    std::vector<void*> owned;
    template <typename T>
    T* upload(const std::vector<T>& h) {
        void* d = nullptr;
        size_t bytes = sizeof(T) * (h.empty() ? 1 : h.size());
        gipc::check(sgipc_malloc(ctx, &d, bytes), ctx, "sgipc_malloc");
        owned.push_back(d);
        if (!h.empty()) gipc::check(sgipc_memcpy_h2d(ctx, d, h.data(), bytes), ctx, "sgipc_memcpy_h2d");
        return static_cast<T*>(d);
    }
    ~DeviceArena() {
        for (void* p : owned) sgipc_free(ctx, p);
    }
};

}  // namespace

// ---- simulator-side stand-ins ---------------------------------------------
// FEM state (the reference's device_TetraData)
struct device_TetraData {
    double* vertexes = nullptr;  // all vertices: body vertices, then FEM
    double* rest_vertexes = nullptr;
    double* xTilde = nullptr;    // FEM part
    double* masses = nullptr;    // FEM part
    int* BoundaryType = nullptr; // all vertices
    uint32_t* tetrahedras = nullptr;
    double* DmInverses = nullptr;
    double* volum = nullptr;
    double* lengthRate = nullptr;  // mu per tet
    double* volumeRate = nullptr;  // lambda per tet
    double* fb = nullptr;          // barrier gradient, all vertices
    double* shape_grads = nullptr; // FEM part
    int tetrahedraNum = 0;
};

struct ABDFemCountInfo {
    int abd_body_num = 0, abd_point_num = 0, fem_point_offset = 0, fem_point_num = 0;
};

namespace gipc {
// affine-body data (the reference's ABDSimData): host body layout
struct ABDSimData {
    std::vector<int> vstart, fixed;
    std::vector<double> xbar, mass12, volume;
    double* q = nullptr;
    double* q_tilde = nullptr;
    double* dq = nullptr;         // solution of the body rows
    double* gradient = nullptr;   // body gradient (12 per body)
};
// affine-body system (the reference's ABDSystem): parameters, set-up
struct ABDSystem {
    struct {
        double kappa = 1e8, dt = 0.01, motor_speed = 0, motor_strength = 0;
    } parms;
    void init_system(GlobalLinearSystem& sys, ABDSimData& d) {
        const int nb = (int)d.volume.size();
        check(sgipc_abd_setup(sys.handle(), nb, d.vstart.back(), d.vstart.data(), d.xbar.data(), d.mass12.data(),
                              d.volume.data(), d.fixed.data()),
              sys.handle(), "sgipc_abd_setup");
    }
};
}  // namespace gipc

struct PCG_Data {
    MASPreconditioner MP;
    int P_type = 1;  // 1: MAS, otherwise 3x3 block-Jacobi
};

class GIPC {
public:
    int vertexNum = 0;
    double* _moveDir = nullptr;
    double pcg_threshold = 1e-4;
    uint32_t h_cpNum[5] = {0, 0, 0, 0, 0};
    ABDFemCountInfo abd_fem_count_info;
    PCG_Data pcg_data;
    gipc::GIPCTripletMatrix gipc_global_triplet;
    std::unique_ptr<gipc::GlobalLinearSystem> m_global_linear_system;
    std::unique_ptr<gipc::ABDSystem> m_abd_system;
    std::unique_ptr<gipc::ABDSimData> m_abd_sim_data;
    gipc::PCGSolver* m_pcg = nullptr;  // kept for its statistics

    void create_LinearSystem(device_TetraData& tet);
};

namespace gipc {
// glue of the FEM vertices (the reference's FEMLinearSubsystem,
// linear_system/subsystem/fem_linear_subsystem.cu:56-102)
class FEMLinearSubsystem : public DiagonalSubsystem {
public:
    FEMLinearSubsystem(::GIPC& gipc, device_TetraData& tet) : m_gipc(gipc), m_tet(tet) {}
    void report_subsystem_info() override { right_hand_side_dof(3 * (SizeT)m_gipc.abd_fem_count_info.fem_point_num); }
    void assemble(DenseVectorView gradient) override {
        const auto& c = m_gipc.abd_fem_count_info;
        check(sgipc_fem_rhs(ctx(), c.fem_point_num, m_tet.fb + 3 * (size_t)c.fem_point_offset, m_tet.shape_grads,
                            m_tet.BoundaryType + c.fem_point_offset, gradient.data()),
              ctx(), "sgipc_fem_rhs");
    }
    void retrieve_solution(CDenseVectorView dx) override {
        const auto& c = m_gipc.abd_fem_count_info;
        check(sgipc_memcpy_d2d(ctx(), m_gipc._moveDir + 3 * (size_t)c.fem_point_offset, dx.data(),
                               sizeof(Float) * dx.size()),
              ctx(), "sgipc_memcpy_d2d");
    }

private:
    ::GIPC& m_gipc;
    device_TetraData& m_tet;
};

// glue of the affine bodies (the reference's ABDLinearSubsystem,
// linear_system/subsystem/abd_linear_subsystem.cu:15-56)
class ABDLinearSubsystem : public DiagonalSubsystem {
public:
    ABDLinearSubsystem(::GIPC& gipc, ABDSystem& abd, ABDSimData& data) : m_gipc(gipc), m_abd(abd), m_data(data) {}
    void report_subsystem_info() override { right_hand_side_dof(12 * (SizeT)m_gipc.abd_fem_count_info.abd_body_num); }
    void assemble(DenseVectorView gradient) override {
        check(sgipc_memcpy_d2d(ctx(), gradient.data(), m_data.gradient, sizeof(Float) * gradient.size()), ctx(),
              "sgipc_memcpy_d2d");
    }
    void retrieve_solution(CDenseVectorView dx) override {
        check(sgipc_memcpy_d2d(ctx(), m_data.dq, dx.data(), sizeof(Float) * dx.size()), ctx(), "sgipc_memcpy_d2d");
    }

private:
    ::GIPC& m_gipc;
    ABDSystem& m_abd;
    ABDSimData& m_data;
};
}  // namespace gipc

// The linear-system set-up: the create<T>() calls of gipc.cu:42-70.
void GIPC::create_LinearSystem(device_TetraData& tet) {
    std::cout << "    - create ABD Linear Subsystem ..." << std::endl;
    auto& abd = m_global_linear_system->create<gipc::ABDLinearSubsystem>(*this, *m_abd_system, *m_abd_sim_data);
    std::cout << "    - create FEM Linear Subsystem ..." << std::endl;
    auto& fem = m_global_linear_system->create<gipc::FEMLinearSubsystem>(*this, tet);

    std::cout << "- create PCG Solver" << std::endl;
    gipc::PCGSolverConfig cfg;
    cfg.global_tol_rate = pcg_threshold;
    auto& pcg = m_global_linear_system->create<gipc::PCGSolver>(cfg);
    m_pcg = &pcg;

    std::cout << "- create Preconditioner" << std::endl;
    m_global_linear_system->create<gipc::ABDPreconditioner>(abd, *m_abd_system, *m_abd_sim_data);
    if (pcg_data.P_type == 1)
        m_global_linear_system->create<gipc::MAS_Preconditioner>(fem, pcg_data.MP, tet.masses, h_cpNum);
    else
        m_global_linear_system->create<gipc::DiagPreconditioner>();
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s scene.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    Reader rd{f};
    Scene s;
    try {
        char magic[8];
        if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "SGIPCHY1", 8) != 0)
            throw std::runtime_error("not a hybrid scene file");
        s.nb = rd.one<int>(); s.nav = rd.one<int>(); s.nf = rd.one<int>(); s.T = rd.one<int>();
        s.bank = rd.one<int>(); s.npairs = rd.one<int>();
        for (int k = 0; k < 3; ++k) s.counts[k] = rd.one<int>();
        (void)rd.one<int>();  // stop rule: the reference rule applies
        (void)rd.one<int>();  // iteration cap: the reference ratio applies
        s.dt = rd.one<double>(); s.dhat = rd.one<double>(); s.kappa = rd.one<double>();
        s.kappa_abd = rd.one<double>(); s.tol = rd.one<double>();
        if (s.nb <= 0 || s.nav <= 0 || s.nf <= 0 || s.T <= 0 || s.npairs < 0) throw std::runtime_error("empty scene");
        const size_t nv = (size_t)s.nav + s.nf;
        rd.get(s.vstart, s.nb + 1); rd.get(s.xbar, 3 * (size_t)s.nav); rd.get(s.mass12, 144 * (size_t)s.nb);
        rd.get(s.volume, s.nb); rd.get(s.fixed, s.nb); rd.get(s.q, 12 * (size_t)s.nb); rd.get(s.qt, 12 * (size_t)s.nb);
        rd.get(s.x, 3 * nv); rd.get(s.xrest, 3 * nv); rd.get(s.btype, nv); rd.get(s.xt, 3 * (size_t)s.nf);
        rd.get(s.mass, s.nf); rd.get(s.tets, 4 * (size_t)s.T); rd.get(s.DmInv, 9 * (size_t)s.T); rd.get(s.vol, s.T);
        rd.get(s.mu, s.T); rd.get(s.lam, s.T); rd.get(s.part, s.nf); rd.get(s.pairs, 4 * (size_t)s.npairs);
        rd.get(s.mat_index, s.npairs);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "bad scene: %s\n", e.what());
        std::fclose(f);
        return 2;
    }
    std::fclose(f);

    int rc = 0;
    try {
        GIPC ipc;
        ipc.m_global_linear_system = std::make_unique<gipc::GlobalLinearSystem>(0);
        gipc::GlobalLinearSystem& sys = *ipc.m_global_linear_system;
        DeviceArena mem{sys.handle(), {}};
        const size_t nv = (size_t)s.nav + s.nf;
        ipc.vertexNum = (int)nv;
        ipc.pcg_threshold = s.tol;
        ipc.abd_fem_count_info = ABDFemCountInfo{s.nb, s.nav, s.nav, s.nf};
        ipc.h_cpNum[0] = (uint32_t)s.npairs;

        device_TetraData tet;
        tet.vertexes = mem.upload(s.x);
        tet.rest_vertexes = mem.upload(s.xrest);
        tet.xTilde = mem.upload(s.xt);
        tet.masses = mem.upload(s.mass);
        tet.BoundaryType = mem.upload(s.btype);
        tet.tetrahedras = mem.upload(s.tets);
        tet.DmInverses = mem.upload(s.DmInv);
        tet.volum = mem.upload(s.vol);
        tet.lengthRate = mem.upload(s.mu);
        tet.volumeRate = mem.upload(s.lam);
        tet.fb = mem.upload(std::vector<double>(3 * nv, 0.0));
        tet.shape_grads = mem.upload(std::vector<double>(3 * (size_t)s.nf, 0.0));
        tet.tetrahedraNum = s.T;
        ipc._moveDir = mem.upload(std::vector<double>(3 * nv, 0.0));
        int* d_pairs = mem.upload(s.pairs);
        int* d_mat_index = mem.upload(s.mat_index);

        ipc.m_abd_system = std::make_unique<gipc::ABDSystem>();
        ipc.m_abd_system->parms.kappa = s.kappa_abd;
        ipc.m_abd_system->parms.dt = s.dt;
        ipc.m_abd_sim_data = std::make_unique<gipc::ABDSimData>();
        gipc::ABDSimData& ad = *ipc.m_abd_sim_data;
        ad.vstart = s.vstart; ad.fixed = s.fixed; ad.xbar = s.xbar; ad.mass12 = s.mass12; ad.volume = s.volume;
        ad.q = mem.upload(s.q);
        ad.q_tilde = mem.upload(s.qt);
        ad.dq = mem.upload(std::vector<double>(12 * (size_t)s.nb, 0.0));
        ad.gradient = mem.upload(std::vector<double>(12 * (size_t)s.nb, 0.0));

        // one-time set-up: triplet store, bodies, FEM mesh, MAS partition
        const long long m_src = 16LL * s.counts[2] + 9LL * s.counts[1] + 4LL * s.counts[0];
        const long long cap = 17 * m_src + 10LL * s.T + s.nf + 10LL * s.nb;
        ipc.gipc_global_triplet.bind(mem.upload(std::vector<double>(9 * (size_t)cap, 0.0)),
                                     mem.upload(std::vector<int>((size_t)cap, 0)),
                                     mem.upload(std::vector<int>((size_t)cap, 0)), cap);
        sys.gipc_global_triplet = &ipc.gipc_global_triplet;
        ipc.m_abd_system->init_system(sys, ad);
        gipc::check(sgipc_set_fem_mesh(sys.handle(), s.nf, s.T, s.tets.data()), sys.handle(), "sgipc_set_fem_mesh");
        ipc.pcg_data.MP.init_partition(s.nf, s.T, s.tets.data(), s.part.data(), s.bank, s.nav, d_pairs);
        ipc.create_LinearSystem(tet);

        // GIPC::computeGradientAndHessian
        gipc::Assembler as(sys, s.nb);
        as.begin();
        as.body_positions(ad.q, tet.vertexes);
        as.barrier_with_bodies(s.npairs, d_pairs, d_mat_index, s.counts, tet.vertexes, tet.rest_vertexes, s.dhat,
                               s.kappa, tet.BoundaryType, tet.fb);
        double* x_fem = tet.vertexes + 3 * (size_t)s.nav;
        as.kinetic_gradient(s.nf, x_fem, tet.xTilde, tet.masses, tet.shape_grads);
        as.fem(s.T, x_fem, tet.tetrahedras, tet.DmInverses, tet.volum, tet.lengthRate, tet.volumeRate, s.dt,
               tet.BoundaryType + s.nav, tet.shape_grads);
        as.mass_diagonal(s.nf, tet.masses);
        as.bodies(ad.q, ad.q_tilde, ipc.m_abd_system->parms.kappa, ipc.m_abd_system->parms.dt, tet.fb, ad.gradient);
        // GIPC::calculateMovingDirection
        const gipc::SizeT iters = sys.solve_linear_system();

        std::vector<double> dx(12 * (size_t)s.nb + 3 * (size_t)s.nf);
        gipc::check(sgipc_memcpy_d2h(sys.handle(), dx.data(), ad.dq, sizeof(double) * 12 * s.nb), sys.handle(),
                    "sgipc_memcpy_d2h");
        gipc::check(sgipc_memcpy_d2h(sys.handle(), dx.data() + 12 * (size_t)s.nb, ipc._moveDir + 3 * (size_t)s.nav,
                                     sizeof(double) * 3 * s.nf),
                    sys.handle(), "sgipc_memcpy_d2h");
        FILE* o = std::fopen(argv[2], "wb");
        if (!o) throw std::runtime_error("cannot write output");
        int head[2] = {(int)iters, ipc.m_pcg->last_stats().converged};
        double rr = ipc.m_pcg->last_stats().rel_residual;
        bool ok = std::fwrite(head, sizeof(int), 2, o) == 2 && std::fwrite(&rr, sizeof(double), 1, o) == 1 &&
                  std::fwrite(dx.data(), sizeof(double), dx.size(), o) == dx.size();
        std::fclose(o);
        if (!ok) throw std::runtime_error("short write");
        std::printf("hybrid_step: bodies=%d fem=%d pairs=%d K=%d iterations=%zu\n", s.nb, s.nf, s.npairs,
                    ipc.gipc_global_triplet.h_unique_key_number, iters);
        // buffers are freed (arena) before the system's context is destroyed
    } catch (const gipc::Error& e) {
        std::fprintf(stderr, "sgipc error %d: %s\n", e.code(), e.what());
        rc = 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        rc = 3;
    }
    return rc;
}
