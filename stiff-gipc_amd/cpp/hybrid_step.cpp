// hybrid_step -- one Newton step of a hybrid scene (affine bodies over a FEM
// body) through the C++ facade (include/sgipc_linear_system.hpp), in the
// order GIPC's Newton loop assembles the coupled system: contact barrier in
// vertex space mapped to DOF space, FEM kinetic and SNK1 terms, mass
// diagonal, body terms with the contact forces of the body vertices, the
// FEM right-hand side, then solve_linear_system() (MAS on the FEM rows,
// 12x12 block-Jacobi on the body rows).
//
// usage: hybrid_step <scene.bin> <out.bin>
// scene.bin (little endian): "SGIPCHY1", int32 nb, nav, nf, T, bank, npairs,
// counts[3], stop_rule, max_iter, double dt, dhat, kappa, kappa_abd, tol,
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
    int nb, nav, nf, T, bank, npairs, counts[3], stop_rule, max_iter;
    double dt, dhat, kappa, kappa_abd, tol;
    std::vector<int> vstart, fixed, btype, part, pairs, mat_index;
    std::vector<double> xbar, mass12, volume, q, qt, x, xrest, xt, mass, DmInv, vol, mu, lam;
    std::vector<uint32_t> tets;
    try {
        char magic[8];
        if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "SGIPCHY1", 8) != 0)
            throw std::runtime_error("not a hybrid scene file");
        nb = rd.one<int>(); nav = rd.one<int>(); nf = rd.one<int>(); T = rd.one<int>(); bank = rd.one<int>();
        npairs = rd.one<int>();
        for (int k = 0; k < 3; ++k) counts[k] = rd.one<int>();
        stop_rule = rd.one<int>(); max_iter = rd.one<int>();
        dt = rd.one<double>(); dhat = rd.one<double>(); kappa = rd.one<double>(); kappa_abd = rd.one<double>();
        tol = rd.one<double>();
        if (nb <= 0 || nav <= 0 || nf <= 0 || T <= 0 || npairs < 0) throw std::runtime_error("empty scene");
        const size_t nv = (size_t)nav + nf;
        rd.get(vstart, nb + 1); rd.get(xbar, 3 * (size_t)nav); rd.get(mass12, 144 * (size_t)nb); rd.get(volume, nb);
        rd.get(fixed, nb); rd.get(q, 12 * (size_t)nb); rd.get(qt, 12 * (size_t)nb); rd.get(x, 3 * nv);
        rd.get(xrest, 3 * nv); rd.get(btype, nv); rd.get(xt, 3 * (size_t)nf); rd.get(mass, nf);
        rd.get(tets, 4 * (size_t)T); rd.get(DmInv, 9 * (size_t)T); rd.get(vol, T); rd.get(mu, T); rd.get(lam, T);
        rd.get(part, nf); rd.get(pairs, 4 * (size_t)npairs); rd.get(mat_index, npairs);
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
            // one-time setup: bodies first (the MAS topology then knows the
            // FEM vertices' offset in vertex space)
            sys.set_affine_bodies(nb, nav, vstart.data(), xbar.data(), mass12.data(), volume.data(), fixed.data());
            sys.set_fem_mesh(nf, T, tets.data());
            sys.set_mas_partition(nf, T, tets.data(), part.data(), bank);
            sys.set_preconditioner(SGIPC_PRECOND_MAS);
            sys.config().global_tol_rate = tol;
            sys.config().stop_rule = stop_rule;
            sys.config().max_iter = max_iter;
            const long long m_src = 16LL * counts[2] + 9LL * counts[1] + 4LL * counts[0];
            const long long cap = 17 * m_src + 10LL * T + nf + 10LL * nb;
            double* d_vals = upload(sys, std::vector<double>(9 * (size_t)cap, 0.0), owned);
            int* d_rows = upload(sys, std::vector<int>((size_t)cap, 0), owned);
            int* d_cols = upload(sys, std::vector<int>((size_t)cap, 0), owned);
            sys.set_triplet_matrix(d_vals, d_rows, d_cols, cap);
            double* d_q = upload(sys, q, owned);
            double* d_qt = upload(sys, qt, owned);
            double* d_x = upload(sys, x, owned);
            double* d_xrest = upload(sys, xrest, owned);
            int* d_btype = upload(sys, btype, owned);
            double* d_xt = upload(sys, xt, owned);
            double* d_mass = upload(sys, mass, owned);
            uint32_t* d_tets = upload(sys, tets, owned);
            double* d_Dm = upload(sys, DmInv, owned);
            double* d_vol = upload(sys, vol, owned);
            double* d_mu = upload(sys, mu, owned);
            double* d_lam = upload(sys, lam, owned);
            int* d_pairs = upload(sys, pairs, owned);
            int* d_mi = upload(sys, mat_index, owned);
            const size_t nrow3 = 12 * (size_t)nb + 3 * (size_t)nf;
            double* d_fb = upload(sys, std::vector<double>(3 * ((size_t)nav + nf), 0.0), owned);
            double* d_shape = upload(sys, std::vector<double>(3 * (size_t)nf, 0.0), owned);
            double* d_b = upload(sys, std::vector<double>(nrow3, 0.0), owned);
            double* d_dx = upload(sys, std::vector<double>(nrow3, 0.0), owned);
            double* d_xf = d_x + 3 * (size_t)nav;

            sys.begin_assembly();
            sys.body_positions(d_q, d_x);
            sys.assemble_barrier_with_bodies(npairs, d_pairs, d_mi, counts, d_x, d_xrest, dhat, kappa, d_btype, d_fb);
            sys.kinetic_gradient(nf, d_xf, d_xt, d_mass, d_shape);
            sys.assemble_fem(T, d_xf, d_tets, d_Dm, d_vol, d_mu, d_lam, dt, d_btype + nav, d_shape);
            sys.mass_diagonal(nf, d_mass);
            sys.assemble_bodies(d_q, d_qt, kappa_abd, dt, d_fb, d_b);
            sys.assemble_rhs(nf, d_fb + 3 * (size_t)nav, d_shape, d_btype + nav, d_b + 12 * (size_t)nb);
            sys.set_contact_pairs(d_pairs, npairs);
            const std::size_t iters = sys.solve_linear_system(d_b, d_dx);
            std::vector<double> dx(nrow3);
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
            std::printf("hybrid_step: bodies=%d fem=%d pairs=%d iterations=%zu rel_residual=%.3e\n", nb, nf, npairs,
                        iters, rr);
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
