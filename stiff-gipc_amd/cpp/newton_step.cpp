// newton_step -- one Newton step of a FEM scene through the C++ facade
// (include/sgipc_linear_system.hpp), driving the library the way GIPC's
// Newton loop drives its linear system: kinetic gradient, SNK1 assembly,
// mass diagonal, right-hand side, then solve_linear_system().
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

#include "sgipc_linear_system.hpp"

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

// device copy of a host vector, freed with the context
template <typename T>
T* upload(sgipc::GlobalLinearSystem& sys, const std::vector<T>& h, std::vector<void*>& owned) {
    void* d = nullptr;
    size_t bytes = sizeof(T) * (h.empty() ? 1 : h.size());
    sgipc::check(sgipc_malloc(sys.handle(), &d, bytes), sys.handle(), "sgipc_malloc");
    owned.push_back(d);
    if (!h.empty()) sgipc::check(sgipc_memcpy_h2d(sys.handle(), d, h.data(), bytes), sys.handle(), "sgipc_memcpy_h2d");
    return static_cast<T*>(d);
}

}  // namespace

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
        rd.get(x, 3 * (size_t)n); rd.get(xt, 3 * (size_t)n); rd.get(mass, n); rd.get(btype, n);
        rd.get(tets, 4 * (size_t)T); rd.get(DmInv, 9 * (size_t)T); rd.get(vol, T); rd.get(mu, T); rd.get(lam, T);
        rd.get(part, n);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "bad scene: %s\n", e.what());
        std::fclose(f);
        return 2;
    }
    std::fclose(f);

    std::vector<void*> owned;
    int rc = 0;
    try {
        sgipc::GlobalLinearSystem sys(0);
        try {
            sys.set_fem_mesh(n, T, tets.data());
            sys.set_mas_partition(n, T, tets.data(), part.data(), bank);
            sys.set_preconditioner(precond);
            sys.config().global_tol_rate = tol;
            sys.config().stop_rule = stop_rule;
            sys.config().max_iter = max_iter;
            const long long cap = 10LL * T + n;
            double* d_vals = upload(sys, std::vector<double>(9 * (size_t)cap, 0.0), owned);
            int* d_rows = upload(sys, std::vector<int>((size_t)cap, 0), owned);
            int* d_cols = upload(sys, std::vector<int>((size_t)cap, 0), owned);
            sys.set_triplet_matrix(d_vals, d_rows, d_cols, cap);
            double* d_x = upload(sys, x, owned);
            double* d_xt = upload(sys, xt, owned);
            double* d_mass = upload(sys, mass, owned);
            int* d_btype = upload(sys, btype, owned);
            uint32_t* d_tets = upload(sys, tets, owned);
            double* d_Dm = upload(sys, DmInv, owned);
            double* d_vol = upload(sys, vol, owned);
            double* d_mu = upload(sys, mu, owned);
            double* d_lam = upload(sys, lam, owned);
            std::vector<double> zeros(3 * (size_t)n, 0.0);
            double* d_shape = upload(sys, zeros, owned);
            double* d_b = upload(sys, zeros, owned);
            double* d_dx = upload(sys, zeros, owned);

            sys.begin_assembly();
            sys.kinetic_gradient(n, d_x, d_xt, d_mass, d_shape);
            sys.assemble_fem(T, d_x, d_tets, d_Dm, d_vol, d_mu, d_lam, dt, d_btype, d_shape);
            sys.mass_diagonal(n, d_mass);
            sys.assemble_rhs(n, nullptr, d_shape, d_btype, d_b);
            const std::size_t iters = sys.solve_linear_system(d_b, d_dx);
            std::vector<double> dx(3 * (size_t)n);
            sgipc::check(sgipc_memcpy_d2h(sys.handle(), dx.data(), d_dx, sizeof(double) * dx.size()), sys.handle(),
                         "sgipc_memcpy_d2h");
            FILE* o = std::fopen(argv[2], "wb");
            if (!o) throw std::runtime_error("cannot write output");
            int head[2] = {(int)iters, sys.last_stats().converged};
            double rr = sys.last_stats().rel_residual;
            bool ok = std::fwrite(head, sizeof(int), 2, o) == 2 && std::fwrite(&rr, sizeof(double), 1, o) == 1 &&
                      std::fwrite(dx.data(), sizeof(double), dx.size(), o) == dx.size();
            std::fclose(o);
            if (!ok) throw std::runtime_error("short write");
            std::printf("newton_step: n=%d T=%d iterations=%zu rel_residual=%.3e\n", n, T, iters, rr);
        } catch (...) {
            for (void* p : owned) sgipc_free(sys.handle(), p);
            throw;
        }
        for (void* p : owned) sgipc_free(sys.handle(), p);
    } catch (const sgipc::Error& e) {
        std::fprintf(stderr, "sgipc error %d: %s\n", e.code(), e.what());
        rc = 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        rc = 3;
    }
    return rc;
}
