// newton_step -- one Newton step of a FEM scene through the reference-shaped
// C++ API (include/gipc_linear_system.hpp): an FEM subsystem, the PCG
// solver and the configured preconditioner registered with
// GlobalLinearSystem::create<T>(), the assembly written into the caller's
// triplet store (kinetic gradient, SNK1 terms, mass diagonal), then
// solve_linear_system(), which assembles the right-hand side through the
// subsystem, converts, builds the preconditioner, solves and hands the
// subsystem its solution.
//
// usage: newton_step <scene.bin> <out.bin>
// scene.bin (little endian): "SGIPCSC1", int32 n, T, bank, precond, stop_rule,
// max_iter, double dt, tol, then x[3n], x_tilde[3n], mass[n] (double),
// btype[n] (int32), tets[4T] (uint32), DmInv[9T], vol[T], mu[T], lam[T]
// (double), part[n] (int32).
// out.bin: int32 iterations, converged, double rel_residual, dx[3n].
// Exit codes: 0 ok, 2 bad input, 3 library error (message on stderr).
#include <cstdio>
#include <cstring>
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

// device buffers freed with the context
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

// the FEM state of the simulator (device arrays)
struct FemState {
    int n = 0;
    const double* fb = nullptr;  // barrier gradient (none in this scene)
    const double* shape = nullptr;
    const int* btype = nullptr;
    double* move_dir = nullptr;
};

}  // namespace

namespace gipc {
// The simulator-side subsystem (the reference's
// linear_system/subsystem/fem_linear_subsystem.cu:56-102) over the library.
class FEMLinearSubsystem : public DiagonalSubsystem {
public:
    explicit FEMLinearSubsystem(FemState& s) : m_s(s) {}
    void report_subsystem_info() override { right_hand_side_dof(3 * (SizeT)m_s.n); }
    void assemble(DenseVectorView gradient) override {
        check(sgipc_fem_rhs(ctx(), m_s.n, m_s.fb, m_s.shape, m_s.btype, gradient.data()), ctx(), "sgipc_fem_rhs");
    }
    void retrieve_solution(CDenseVectorView dx) override {
        check(sgipc_memcpy_d2d(ctx(), m_s.move_dir, dx.data(), sizeof(Float) * dx.size()), ctx(), "sgipc_memcpy_d2d");
    }

private:
    FemState& m_s;
};
}  // namespace gipc

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
    int n, T, bank, precond, stop_rule, max_iter;
    double dt, tol;
    std::vector<double> x, xt, mass, DmInv, vol, mu, lam;
    std::vector<int> btype, part;
    std::vector<uint32_t> tets;
    try {
        char magic[8];
        if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "SGIPCSC1", 8) != 0)
            throw std::runtime_error("not a scene file");
        n = rd.one<int>(); T = rd.one<int>(); bank = rd.one<int>();
        precond = rd.one<int>(); stop_rule = rd.one<int>(); max_iter = rd.one<int>();
        dt = rd.one<double>(); tol = rd.one<double>();
        if (n <= 0 || T <= 0) throw std::runtime_error("empty scene");
        if (precond < SGIPC_PRECOND_NONE || precond > SGIPC_PRECOND_MAS_FP32) throw std::runtime_error("bad preconditioner");
        rd.get(x, 3 * (size_t)n); rd.get(xt, 3 * (size_t)n); rd.get(mass, n); rd.get(btype, n);
        rd.get(tets, 4 * (size_t)T); rd.get(DmInv, 9 * (size_t)T); rd.get(vol, T); rd.get(mu, T); rd.get(lam, T);
        rd.get(part, n);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "bad scene: %s\n", e.what());
        std::fclose(f);
        return 2;
    }
    std::fclose(f);

    int rc = 0;
    try {
        gipc::GlobalLinearSystem sys(0);
        DeviceArena mem{sys.handle(), {}};
        gipc::check(sgipc_set_fem_mesh(sys.handle(), n, T, tets.data()), sys.handle(), "sgipc_set_fem_mesh");
        MASPreconditioner mp;
        mp.init_partition(n, T, tets.data(), part.data(), bank);
        uint32_t cp_num[5] = {0, 0, 0, 0, 0};

        const long long cap = 10LL * T + n;
        gipc::GIPCTripletMatrix trip(mem.upload(std::vector<double>(9 * (size_t)cap, 0.0)),
                                     mem.upload(std::vector<int>((size_t)cap, 0)),
                                     mem.upload(std::vector<int>((size_t)cap, 0)), cap);
        sys.gipc_global_triplet = &trip;
        double* d_x = mem.upload(x);
        double* d_xt = mem.upload(xt);
        double* d_mass = mem.upload(mass);
        int* d_btype = mem.upload(btype);
        uint32_t* d_tets = mem.upload(tets);
        double* d_Dm = mem.upload(DmInv);
        double* d_vol = mem.upload(vol);
        double* d_mu = mem.upload(mu);
        double* d_lam = mem.upload(lam);
        std::vector<double> zeros(3 * (size_t)n, 0.0);
        FemState fs;
        fs.n = n;
        double* d_shape = mem.upload(zeros);
        fs.shape = d_shape;
        fs.btype = d_btype;
        fs.move_dir = mem.upload(zeros);

        // linear system set-up, as GIPC::create_LinearSystem does it
        auto& fem = sys.create<gipc::FEMLinearSubsystem>(fs);
        gipc::PCGSolverConfig cfg;
        cfg.global_tol_rate = tol;
        cfg.stop_rule = stop_rule;
        cfg.max_iter = max_iter;
        auto& pcg = sys.create<gipc::PCGSolver>(cfg);
        if (precond == SGIPC_PRECOND_MAS || precond == SGIPC_PRECOND_MAS_FP32)
            sys.create<gipc::MAS_Preconditioner>(fem, mp, d_mass, cp_num, precond);
        else if (precond == SGIPC_PRECOND_DIAG)
            sys.create<gipc::DiagPreconditioner>();

        // GIPC::computeGradientAndHessian
        gipc::Assembler as(sys);
        as.begin();
        as.kinetic_gradient(n, d_x, d_xt, d_mass, d_shape);
        as.fem(T, d_x, d_tets, d_Dm, d_vol, d_mu, d_lam, dt, d_btype, d_shape);
        as.mass_diagonal(n, d_mass);
        // GIPC::calculateMovingDirection
        const gipc::SizeT iters = sys.solve_linear_system();

        std::vector<double> dx(3 * (size_t)n);
        gipc::check(sgipc_memcpy_d2h(sys.handle(), dx.data(), fs.move_dir, sizeof(double) * dx.size()), sys.handle(),
              "sgipc_memcpy_d2h");
        FILE* o = std::fopen(argv[2], "wb");
        if (!o) throw std::runtime_error("cannot write output");
        int head[2] = {(int)iters, pcg.last_stats().converged};
        double rr = pcg.last_stats().rel_residual;
        bool ok = std::fwrite(head, sizeof(int), 2, o) == 2 && std::fwrite(&rr, sizeof(double), 1, o) == 1 &&
                  std::fwrite(dx.data(), sizeof(double), dx.size(), o) == dx.size();
        std::fclose(o);
        if (!ok) throw std::runtime_error("short write");
        std::printf("newton_step: n=%d T=%d K=%d iterations=%zu rel_residual=%.3e\n", n, T, trip.h_unique_key_number,
                    iters, rr);
    } catch (const gipc::Error& e) {
        std::fprintf(stderr, "sgipc error %d: %s\n", e.code(), e.what());
        rc = 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        rc = 3;
    }
    return rc;
}
