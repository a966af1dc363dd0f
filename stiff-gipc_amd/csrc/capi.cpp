// extern "C" boundary of libsgipc (declarations: include/sgipc.h).
#include <cstdlib>
#include <cstring>
#include <vector>
#include "../../include/sgipc.h"
#include "context.hpp"

#include <new>

namespace sgipc {
int build_vertex_neighbors(int n, int T, const uint32_t* tets, uint32_t* start, uint32_t* num, uint32_t* list,
                           long long cap);
int build_bank_map(int n, const int* part, int bank, int* part_map, long long cap, int* real_map);
int partition_mesh(int n, int T, const uint32_t* tets, int bank, int* part, int* order, int* offset);
long long partition_edge_cut(int n, int T, const uint32_t* tets, const int* part);
int mas_num_levels(int nn, int bank);
}  // namespace sgipc

struct sgipc_ctx : sgipc::Ctx {};

namespace {
thread_local std::string g_free_error;

template <typename F>
int guard(sgipc_ctx* c, F&& f) {
    try {
        if (c) c->last_error.clear();
        f();
        return SGIPC_OK;
    } catch (const sgipc::arg_error& e) {
        if (c) c->last_error = e.what(); else g_free_error = e.what();
        return SGIPC_ERR_INVALID_ARG;
    } catch (const std::invalid_argument& e) {
        if (c) c->last_error = e.what(); else g_free_error = e.what();
        return SGIPC_ERR_INVALID_ARG;
    } catch (const sgipc::capacity_error& e) {
        if (c) c->last_error = e.what(); else g_free_error = e.what();
        return SGIPC_ERR_CAPACITY;
    } catch (const std::length_error& e) {
        if (c) c->last_error = e.what(); else g_free_error = e.what();
        return SGIPC_ERR_CAPACITY;
    } catch (const sgipc::state_error& e) {
        if (c) c->last_error = e.what(); else g_free_error = e.what();
        return SGIPC_ERR_STATE;
    } catch (const sgipc::hip_error& e) {
        if (c) c->last_error = e.what(); else g_free_error = e.what();
        return SGIPC_ERR_HIP;
    } catch (const std::exception& e) {
        if (c) c->last_error = e.what(); else g_free_error = e.what();
        return SGIPC_ERR_INTERNAL;
    } catch (...) {
        if (c) c->last_error = "unknown error"; else g_free_error = "unknown error";
        return SGIPC_ERR_INTERNAL;
    }
}

void set_device(sgipc_ctx* c) { SGIPC_HIP(hipSetDevice(c->device)); }

#define CHECK_CTX(c)                                                     \
    do {                                                                 \
        if (!(c)) { g_free_error = "null context"; return SGIPC_ERR_INVALID_ARG; } \
    } while (0)
}  // namespace

extern "C" {

int sgipc_version(void) { return 1; }

int sgipc_ctx_create(int device, sgipc_ctx** out) {
    if (!out) { g_free_error = "null output pointer"; return SGIPC_ERR_INVALID_ARG; }
    *out = nullptr;
    sgipc_ctx* c = new (std::nothrow) sgipc_ctx();
    if (!c) { g_free_error = "out of host memory"; return SGIPC_ERR_INTERNAL; }
    int rc = guard(nullptr, [&] {
        int ndev = 0;
        SGIPC_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw sgipc::arg_error("invalid device index");
        c->device = device;
        SGIPC_HIP(hipSetDevice(device));
        // this device's address of the allocation-free energy's scratch
        sgipc::energy_scratch_setup(device);
        // blocking stream: ordered with work on the legacy null stream, which
        // callers (and torch's default stream) may use on the same buffers
        SGIPC_HIP(hipStreamCreate(&c->own_stream));
        c->stream = c->own_stream;
        // mixed-precision PCG SpMV: one resident round of 1024-thread
        // workgroups (two per CU) walking the row chunks; a second, partial
        // round measured slower
        int ncu = 0;
        SGIPC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        if (ncu > 0) c->spmv_blocks = std::min(1024, 2 * ncu);
        if (const char* g = std::getenv("SGIPC_SPMV_BLOCKS")) {
            const int v = std::atoi(g);
            if (v >= 8 && v <= 1024) c->spmv_blocks = v;
        }
        if (const char* g = std::getenv("SGIPC_SPMV_G")) {
            int v = std::atoi(g);
            if (v == 4 || v == 8 || v == 16) c->spmv_group = v;
        }
        if (const char* f = std::getenv("SGIPC_SPMV_FULL32")) c->spmv_full32 = std::atoi(f) != 0;
        if (const char* f = std::getenv("SGIPC_MAS_FUSED")) c->mas_fused = std::atoi(f) != 0;
        if (const char* f = std::getenv("SGIPC_SORT")) c->sort_impl = std::string(f) == "rocprim" ? 1 : 0;
        if (const char* f = std::getenv("SGIPC_INNER_TOL")) {
            const double v = std::atof(f);
            if (v >= 0.0 && v < 1.0) c->inner_tol = v;
        }
        if (const char* f = std::getenv("SGIPC_BANK_INVERSE")) {
            const int v = std::atoi(f);
            if (v >= SGIPC_BANK_INVERSE_VALU64 && v <= SGIPC_BANK_INVERSE_MFMA32_REFINED) c->bank_inverse = v;
        }
    });
    if (rc != SGIPC_OK) { delete c; return rc; }
    *out = c;
    return SGIPC_OK;
}

int sgipc_ctx_destroy(sgipc_ctx* c) {
    if (!c) return SGIPC_OK;
    (void)hipSetDevice(c->device);
    sgipc::prof_release(*c);
    if (c->own_stream) {
        (void)hipStreamSynchronize(c->own_stream);
        (void)hipStreamDestroy(c->own_stream);
    }
    delete c;
    return SGIPC_OK;
}

int sgipc_ctx_set_stream(sgipc_ctx* c, void* stream) {
    CHECK_CTX(c);
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return SGIPC_OK;
}

int sgipc_ctx_synchronize(sgipc_ctx* c) {
    CHECK_CTX(c);
    return guard(c, [&] { set_device(c); SGIPC_HIP(hipStreamSynchronize(c->stream)); });
}

const char* sgipc_last_error(const sgipc_ctx* c) { return c ? c->last_error.c_str() : g_free_error.c_str(); }

int sgipc_build_vertex_neighbors(int n, int T, const uint32_t* tets, uint32_t* start, uint32_t* num, uint32_t* list,
                                 int cap, int* len) {
    return guard(nullptr, [&] {
        if (n < 0 || T < 0 || (T > 0 && !tets) || !start || !num || !list || !len)
            throw sgipc::arg_error("build_vertex_neighbors: invalid argument");
        *len = sgipc::build_vertex_neighbors(n, T, tets, start, num, list, cap);
    });
}

int sgipc_build_bank_map(int n, const int* part, int bank, int* part_map, int cap, int* M, int* real_map) {
    return guard(nullptr, [&] {
        if (n < 0 || (n > 0 && (!part || !part_map || !real_map)) || !M || (bank != 16 && bank != 32))
            throw sgipc::arg_error("build_bank_map: invalid argument");
        *M = sgipc::build_bank_map(n, part, bank, part_map, cap, real_map);
    });
}

int sgipc_mas_num_levels(int nn, int bank) { return sgipc::mas_num_levels(nn, bank); }

int sgipc_partition_mesh(int n, int T, const uint32_t* tets, int bank, int* part, int* order, int* n_parts,
                         int* offset) {
    return guard(nullptr, [&] {
        if (n < 0 || T < 0 || (T > 0 && !tets) || (n > 0 && !part) || !n_parts || (bank != 16 && bank != 32))
            throw sgipc::arg_error("partition_mesh: invalid argument");
        int off = 0;
        *n_parts = sgipc::partition_mesh(n, T, tets, bank, part, order, &off);
        if (offset) *offset = off;
    });
}

int sgipc_partition_edge_cut(int n, int T, const uint32_t* tets, const int* part, long long* cut) {
    return guard(nullptr, [&] {
        if (n < 0 || T < 0 || (T > 0 && !tets) || (n > 0 && !part) || !cut)
            throw sgipc::arg_error("partition_edge_cut: invalid argument");
        *cut = sgipc::partition_edge_cut(n, T, tets, part);
    });
}

int sgipc_set_fem_mesh(sgipc_ctx* c, int n, int T, const uint32_t* h_tets) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (T > 0 && !h_tets) throw sgipc::arg_error("set_fem_mesh: null tets");
        set_device(c);
        sgipc::fem_set_mesh(*c, n, T, h_tets);
    });
}

int sgipc_set_mas_topology(sgipc_ctx* c, int n, int node_offset, const uint32_t* ns, const uint32_t* nn,
                           const uint32_t* nl, int nlen, const int* pm, int M, const int* rm, int bank) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n > 0 && (!ns || !nn || !rm)) throw sgipc::arg_error("set_mas_topology: null array");
        if (M > 0 && !pm) throw sgipc::arg_error("set_mas_topology: null part map");
        if (nlen > 0 && !nl) throw sgipc::arg_error("set_mas_topology: null neighbour list");
        set_device(c);
        sgipc::mas_set_topology(*c, n, node_offset, ns, nn, nl, nlen, pm, M, rm, bank);
    });
}

int sgipc_triplets_view(sgipc_ctx* c, double* v, int* r, int* col, long long cap) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (cap < 0 || (cap > 0 && (!v || !r || !col))) throw sgipc::arg_error("triplets_view: invalid arrays");
        c->t_vals = v;
        c->t_rows = r;
        c->t_cols = col;
        c->t_cap = cap;
        c->have_bcoo = false;
        c->have_bsr = false;
    });
}

int sgipc_kinetic_gradient(sgipc_ctx* c, int n, const double* x, const double* xt, const double* m, double* g) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n < 0 || (n > 0 && (!x || !xt || !m || !g))) throw sgipc::arg_error("kinetic_gradient: invalid argument");
        set_device(c);
        sgipc::fem_kinetic_gradient(*c, n, x, xt, m, g);
    });
}

int sgipc_assemble_fem(sgipc_ctx* c, int T, const double* x, const uint32_t* tets, const double* DmInv,
                       const double* vol, const double* mu, const double* lam, double dt, const int* btype,
                       int index_offset, long long triplet_offset, double* shape) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (T < 0 || (T > 0 && (!x || !tets || !DmInv || !vol || !mu || !lam || !shape)))
            throw sgipc::arg_error("assemble_fem: invalid argument");
        set_device(c);
        sgipc::fem_assemble(*c, T, x, tets, DmInv, vol, mu, lam, dt, btype, index_offset, triplet_offset, shape);
        if (T > 0) c->static_begin = triplet_offset;  // FEM + mass triplets keep their indices across steps
    });
}

int sgipc_mass_triplets(sgipc_ctx* c, int n, const double* m, int index_offset, long long triplet_offset) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n < 0 || (n > 0 && !m)) throw sgipc::arg_error("mass_triplets: invalid argument");
        set_device(c);
        sgipc::fem_mass_triplets(*c, n, m, index_offset, triplet_offset);
    });
}

int sgipc_fem_rhs(sgipc_ctx* c, int n, const double* fb, const double* sh, const int* bt, double* b) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n < 0 || (n > 0 && (!sh || !b))) throw sgipc::arg_error("fem_rhs: invalid argument");
        set_device(c);
        sgipc::fem_rhs(*c, n, fb, sh, bt, b);
    });
}

int sgipc_ground_detect(sgipc_ctx* c, int m, const uint32_t* cand, const double* x, const double normal[3],
                        double offset, double dhat, uint32_t* out, int* n_close) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (m < 0 || !normal || !n_close || (m > 0 && (!cand || !x || !out)))
            throw sgipc::arg_error("ground_detect: invalid argument");
        set_device(c);
        *n_close = sgipc::ground_detect(*c, m, cand, x, normal, offset, dhat, out);
    });
}

int sgipc_assemble_ground(sgipc_ctx* c, int np, const uint32_t* gp, const double* x, const double normal[3],
                          double offset, double dhat, double kappa, const int* bt, int index_offset,
                          long long triplet_offset, double* fb) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (np < 0 || !normal || (np > 0 && (!gp || !x || !fb)) || !(dhat > 0.0))
            throw sgipc::arg_error("assemble_ground: invalid argument");
        set_device(c);
        sgipc::ground_assemble(*c, np, gp, x, normal, offset, dhat, kappa, bt, index_offset, triplet_offset, fb);
    });
}

int sgipc_ground_energy(sgipc_ctx* c, int m, const uint32_t* ids, const double* x, const double normal[3],
                        double offset, double dhat, double kappa, double* e) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (m < 0 || !normal || !e || (m > 0 && (!ids || !x)) || !(dhat > 0.0))
            throw sgipc::arg_error("ground_energy: invalid argument");
        set_device(c);
        *e = sgipc::ground_energy(*c, m, ids, x, normal, offset, dhat, kappa);
    });
}

int sgipc_ground_step_bound(sgipc_ctx* c, int m, const uint32_t* ids, const double* x, const double* dir,
                            const double normal[3], double offset, double slack, double* alpha) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (m < 0 || !normal || !alpha || (m > 0 && (!ids || !x || !dir)))
            throw sgipc::arg_error("ground_step_bound: invalid argument");
        set_device(c);
        *alpha = sgipc::ground_step_bound(*c, m, ids, x, dir, normal, offset, slack);
    });
}

int sgipc_step_forward(sgipc_ctx* c, int n, double* x, const double* x_temp, const double* dir, const int* btype,
                       double alpha, int move_boundary) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n < 0 || (n > 0 && (!x || !x_temp || !dir))) throw sgipc::arg_error("step_forward: invalid argument");
        set_device(c);
        sgipc::step_forward(*c, n, x, x_temp, dir, btype, alpha, move_boundary ? 1 : 0);
    });
}

int sgipc_cfl_largest_speed(sgipc_ctx* c, int n_ids, const uint32_t* ids, const double* dir, double* speed) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n_ids < 0 || !speed || (n_ids > 0 && (!ids || !dir)))
            throw sgipc::arg_error("cfl_largest_speed: invalid argument");
        set_device(c);
        *speed = sgipc::max_speed(*c, n_ids, ids, dir);
    });
}

int sgipc_broad_phase(sgipc_ctx* c, int n_surf_verts, const uint32_t* surf_verts, int n_faces, const uint32_t* faces,
                      int n_edges, const uint32_t* edges, const double* x, const int* btype, const int* body_id,
                      double dhat, int* pt_out, int pt_capacity, int* n_pt, int* ee_out, int ee_capacity, int* n_ee) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n_surf_verts < 0 || n_faces < 0 || n_edges < 0 || pt_capacity < 0 || ee_capacity < 0 || !n_pt || !n_ee ||
            !x || (n_surf_verts > 0 && !surf_verts) || (n_faces > 0 && !faces) || (n_edges > 0 && !edges) ||
            (pt_capacity > 0 && !pt_out) || (ee_capacity > 0 && !ee_out))
            throw sgipc::arg_error("broad_phase: invalid argument");
        set_device(c);
        int a = 0, b = 0;
        sgipc::broad_phase(*c, n_surf_verts, surf_verts, n_faces, faces, n_edges, edges, x, btype, body_id, dhat,
                           pt_out, pt_capacity, a, ee_out, ee_capacity, b);
        *n_pt = a;
        *n_ee = b;
        if (a > pt_capacity || b > ee_capacity)
            throw sgipc::capacity_error("broad_phase: " + std::to_string(a) + " point-triangle / " +
                                        std::to_string(b) + " edge-edge candidates exceed the capacities (" +
                                        std::to_string(pt_capacity) + " / " + std::to_string(ee_capacity) + ")");
    });
}

int sgipc_broad_phase_ccd(sgipc_ctx* c, int n_surf_verts, const uint32_t* surf_verts, int n_faces,
                          const uint32_t* faces, int n_edges, const uint32_t* edges, const double* x, const double* dir,
                          double alpha, const int* btype, const int* body_id, double dhat, int* pt_out,
                          int pt_capacity, int* n_pt, int* ee_out, int ee_capacity, int* n_ee) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!dir || !(alpha >= 0.0)) throw sgipc::arg_error("broad_phase_ccd: a direction and alpha >= 0 are required");
        if (n_surf_verts < 0 || n_faces < 0 || n_edges < 0 || pt_capacity < 0 || ee_capacity < 0 || !n_pt || !n_ee ||
            !x || (n_surf_verts > 0 && !surf_verts) || (n_faces > 0 && !faces) || (n_edges > 0 && !edges) ||
            (pt_capacity > 0 && !pt_out) || (ee_capacity > 0 && !ee_out))
            throw sgipc::arg_error("broad_phase_ccd: invalid argument");
        set_device(c);
        int a = 0, b = 0;
        sgipc::broad_phase(*c, n_surf_verts, surf_verts, n_faces, faces, n_edges, edges, x, btype, body_id, dhat,
                           pt_out, pt_capacity, a, ee_out, ee_capacity, b, dir, alpha);
        *n_pt = a;
        *n_ee = b;
        if (a > pt_capacity || b > ee_capacity)
            throw sgipc::capacity_error("broad_phase_ccd: " + std::to_string(a) + " point-triangle / " +
                                        std::to_string(b) + " edge-edge candidates exceed the capacities (" +
                                        std::to_string(pt_capacity) + " / " + std::to_string(ee_capacity) + ")");
    });
}

int sgipc_ccd_step_bound(sgipc_ctx* c, int n_pt, const int* pt, int n_ee, const int* ee, const double* x,
                         const double* dir, double slackness, double* alpha) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n_pt < 0 || n_ee < 0 || !alpha || !x || !dir || (n_pt > 0 && !pt) || (n_ee > 0 && !ee))
            throw sgipc::arg_error("ccd_step_bound: invalid argument");
        set_device(c);
        *alpha = sgipc::ccd_step_bound(*c, n_pt, pt, n_ee, ee, x, dir, slackness);
    });
}

int sgipc_set_deterministic(sgipc_ctx* c, int on) {
    CHECK_CTX(c);
    return guard(c, [&] { c->deterministic = on ? 1 : 0; });
}

int sgipc_set_fem_projection(sgipc_ctx* c, int on) {
    CHECK_CTX(c);
    return guard(c, [&] { c->fem_project = on ? 1 : 0; });
}

int sgipc_fem_param_adjoint(sgipc_ctx* c, int T, const double* x, const uint32_t* tets, const double* DmInv,
                            const double* vol, double dt, const int* bt, const double* adj, double* dmu,
                            double* dlam) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (T < 0 || (T > 0 && (!x || !tets || !DmInv || !vol || !adj || !dmu || !dlam)))
            throw sgipc::arg_error("fem_param_adjoint: invalid argument");
        set_device(c);
        sgipc::fem_param_adjoint(*c, T, x, tets, DmInv, vol, dt, bt, adj, dmu, dlam);
    });
}

int sgipc_classify_contacts(sgipc_ctx* c, int npt, const int* pt, int nee, const int* ee, const double* x,
                            const double* xr, double dhat, int* pairs, int* mi, int capacity, int* n_pairs,
                            int* counts) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (npt < 0 || nee < 0 || (npt > 0 && !pt) || (nee > 0 && !ee) || !x || !xr || !pairs || !mi || !n_pairs ||
            !counts)
            throw sgipc::arg_error("classify_contacts: invalid argument");
        if (!(dhat > 0.0)) throw sgipc::arg_error("classify_contacts: dhat must be positive");
        set_device(c);
        *n_pairs = sgipc::classify_contacts(*c, npt, pt, nee, ee, x, xr, dhat, pairs, mi, capacity, counts);
    });
}

int sgipc_assemble_barrier(sgipc_ctx* c, int np, const int* pairs, const int* mi, const int* counts, const double* x,
                           const double* xr, double dhat, double kappa, const int* btype, int index_offset,
                           long long triplet_offset, double* fb) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (np < 0 || !counts || (np > 0 && (!pairs || !mi || !x || !xr || !fb)))
            throw sgipc::arg_error("assemble_barrier: invalid argument");
        if (counts[0] < 0 || counts[1] < 0 || counts[2] < 0 || counts[0] + counts[1] + counts[2] != np)
            throw sgipc::arg_error("assemble_barrier: class counts do not add up to the pair count");
        if (!(dhat > 0.0)) throw sgipc::arg_error("assemble_barrier: dhat must be positive");
        set_device(c);
        sgipc::barrier_assemble(*c, np, pairs, mi, counts, x, xr, dhat, kappa, btype, index_offset, triplet_offset, fb);
    });
}

int sgipc_convert(sgipc_ctx* c, long long nt, int nbr, int* n_unique) {
    CHECK_CTX(c);
    return guard(c, [&] {
        set_device(c);
        sgipc::prof_begin(*c, sgipc::PROF_CONVERT);
        int K = sgipc::bcoo_convert(*c, nt, nbr);
        sgipc::prof_end(*c, sgipc::PROF_CONVERT);
        if (n_unique) *n_unique = K;
    });
}

int sgipc_spmv(sgipc_ctx* c, double a, const double* x, double b, double* y) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!x || !y) throw sgipc::arg_error("spmv: null vector");
        if (!c->have_bcoo) throw sgipc::state_error("spmv: convert first");
        set_device(c);
        sgipc::spmv64(*c, a, x, b, y);
    });
}

int sgipc_precond_build(sgipc_ctx* c, int kind, int sub_offset, int sub_nb, const int* pairs, int npairs) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (kind < SGIPC_PRECOND_NONE || kind > SGIPC_PRECOND_MAS_FP32) throw sgipc::arg_error("unknown preconditioner");
        if (npairs < 0 || (npairs > 0 && !pairs)) throw sgipc::arg_error("precond_build: invalid pairs");
        if (!c->have_bcoo) throw sgipc::state_error("precond_build: convert first");
        if (sub_offset < 0 || sub_nb < 0 || sub_offset + sub_nb > c->nb)
            throw sgipc::arg_error("precond_build: subsystem range outside the system");
        set_device(c);
        c->precond_kind = kind;
        c->sub_offset = sub_offset;
        c->sub_nb = sub_nb;
        sgipc::prof_begin(*c, sgipc::PROF_PRECOND_BUILD);
        if (kind == SGIPC_PRECOND_DIAG) sgipc::diag_build(*c);
        if (kind == SGIPC_PRECOND_MAS || kind == SGIPC_PRECOND_MAS_FP32) {
            sgipc::mas_build(*c, sub_offset, sub_nb, pairs, npairs);
            // affine bodies below the MAS range: their 12x12 block-Jacobi
            if (c->abd.ready && sub_offset >= 4 * c->abd.nb) sgipc::abd_jacobi_build(*c);
            else c->abd.inv_ready = false;
        }
        sgipc::prof_end(*c, sgipc::PROF_PRECOND_BUILD);
    });
}

int sgipc_abd_setup(sgipc_ctx* c, int n_bodies, int n_body_vertices, const int* vstart, const double* xbar,
                    const double* mass12, const double* volume, const int* fixed) {
    CHECK_CTX(c);
    return guard(c, [&] {
        set_device(c);
        sgipc::abd_setup(*c, n_bodies, n_body_vertices, vstart, xbar, mass12, volume, fixed);
    });
}

int sgipc_abd_positions(sgipc_ctx* c, const double* q, double* x) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!q || !x) throw sgipc::arg_error("abd_positions: null array");
        set_device(c);
        sgipc::abd_positions(*c, q, x);
    });
}

int sgipc_abd_assemble(sgipc_ctx* c, const double* q, const double* q_tilde, double kappa, double dt,
                       long long triplet_offset, double* grad_q) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!q || !q_tilde || !grad_q) throw sgipc::arg_error("abd_assemble: null array");
        set_device(c);
        sgipc::abd_assemble(*c, q, q_tilde, kappa, dt, triplet_offset, grad_q);
    });
}

int sgipc_abd_contact_gradient(sgipc_ctx* c, const double* vertex_gradient, double* grad_q) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!vertex_gradient || !grad_q) throw sgipc::arg_error("abd_contact_gradient: null array");
        set_device(c);
        sgipc::abd_contact_gradient(*c, vertex_gradient, grad_q);
    });
}

int sgipc_abd_map_contacts(sgipc_ctx* c, long long src_offset, long long count, long long dst_offset,
                           long long* n_written) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!n_written) throw sgipc::arg_error("abd_map_contacts: null output");
        set_device(c);
        *n_written = sgipc::abd_map_contacts(*c, src_offset, count, dst_offset);
    });
}

int sgipc_abd_energy(sgipc_ctx* c, const double* q, const double* q_tilde, double kappa, double dt, double* e) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!q || !q_tilde || !e) throw sgipc::arg_error("abd_energy: null array");
        set_device(c);
        *e = sgipc::abd_energy(*c, q, q_tilde, kappa, dt);
    });
}

int sgipc_precond_apply(sgipc_ctx* c, const double* r, double* z) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!r || !z) throw sgipc::arg_error("precond_apply: null vector");
        set_device(c);
        sgipc::precond_apply(*c, r, z);
    });
}

static sgipc::SolveOpts solve_opts_of(const sgipc_ctx* c, const sgipc_solve_opts* o) {
    sgipc::SolveOpts so;
    so.inner_tol = c->inner_tol;
    if (o) {
        so.precision = o->precision;
        so.stop_rule = o->stop_rule;
        so.tol = o->tol;
        so.max_iter_ratio = o->max_iter_ratio;
        so.max_iter = o->max_iter;
        so.check_every = o->check_every;
    }
    if (so.stop_rule != 0 && so.stop_rule != 1) throw sgipc::arg_error("solve: unknown stop rule");
    if (so.precision != 0 && so.precision != 1) throw sgipc::arg_error("solve: unknown precision");
    if (!(so.tol > 0.0)) throw sgipc::arg_error("solve: tolerance must be positive");
    return so;
}

int sgipc_solve_batched(sgipc_ctx* c, int n_scenes, const int* row_start, const double* b, double* x,
                        const sgipc_solve_opts* o, sgipc_solve_stats* st) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n_scenes <= 0 || !row_start || !b || !x || !st) throw sgipc::arg_error("solve_batched: invalid argument");
        sgipc::SolveOpts so = solve_opts_of(c, o);
        set_device(c);
        std::vector<sgipc::SolveStats> s2(n_scenes);
        sgipc::prof_begin(*c, sgipc::PROF_SOLVE);
        sgipc::pcg_solve_batched(*c, n_scenes, row_start, b, x, so, s2.data());
        sgipc::prof_end(*c, sgipc::PROF_SOLVE);
        sgipc::prof_flush(*c);
        for (int i = 0; i < n_scenes; ++i) {
            st[i].iterations = s2[i].iterations;
            st[i].converged = s2[i].converged;
            st[i].refinements = s2[i].refinements;
            st[i].rel_residual = s2[i].rel_residual;
            st[i].rz0 = s2[i].rz0;
        }
    });
}

int sgipc_autotune(sgipc_ctx* c, const double* b, int n_scenes, const int* row_start, int precision, int* group,
                   int* iters_per_graph) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!b || n_scenes < 0 || (n_scenes > 0 && !row_start) || (precision != 0 && precision != 1))
            throw sgipc::arg_error("autotune: invalid argument");
        set_device(c);
        int g = 0, pg = 0;
        sgipc::autotune(*c, b, n_scenes, row_start, precision, g, pg);
        if (group) *group = g;
        if (iters_per_graph) *iters_per_graph = pg;
    });
}

int sgipc_set_schedule(sgipc_ctx* c, int group, int iters_per_graph) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (group != 4 && group != 8 && group != 16) throw sgipc::arg_error("set_schedule: group must be 4, 8 or 16");
        if (iters_per_graph < 1 || iters_per_graph > 64)
            throw sgipc::arg_error("set_schedule: iterations per graph must be in [1, 64]");
        c->spmv_group = group;
        c->batch_iters_per_graph = iters_per_graph;
    });
}

int sgipc_set_batch_mode(sgipc_ctx* c, int persistent, int threads, int max_workgroups) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (persistent != 0 && persistent != 1) throw sgipc::arg_error("set_batch_mode: persistent must be 0 or 1");
        if (threads != 512 && threads != 1024) throw sgipc::arg_error("set_batch_mode: threads must be 512 or 1024");
        if (max_workgroups < 0) throw sgipc::arg_error("set_batch_mode: max_workgroups must be >= 0");
        c->batch_persistent = persistent;
        c->batch_threads = threads;
        c->batch_max_wg = max_workgroups;
    });
}

int sgipc_set_inner_tol(sgipc_ctx* c, double floor_tol) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!(floor_tol >= 0.0 && floor_tol < 1.0)) throw sgipc::arg_error("set_inner_tol: must be in [0, 1)");
        c->inner_tol = floor_tol;
    });
}

int sgipc_set_sort(sgipc_ctx* c, int impl) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (impl != SGIPC_SORT_OWN && impl != SGIPC_SORT_ROCPRIM) throw sgipc::arg_error("set_sort: unknown sort");
        c->sort_impl = impl;
    });
}

int sgipc_sort_pairs(sgipc_ctx* c, const void* d_keys_in, void* d_keys_out, const int* d_vals_in, int* d_vals_out,
                     long long n, int key_bytes, int bits, int impl) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n < 0 || (n > 0 && (!d_keys_in || !d_keys_out || !d_vals_out)))
            throw sgipc::arg_error("sort_pairs: invalid argument");
        if (key_bytes != 4 && key_bytes != 8) throw sgipc::arg_error("sort_pairs: keys are 4 or 8 bytes");
        if (bits < 1 || bits > 8 * key_bytes) throw sgipc::arg_error("sort_pairs: bits must be in [1, 8 key_bytes]");
        if (impl != SGIPC_SORT_OWN && impl != SGIPC_SORT_ROCPRIM) throw sgipc::arg_error("sort_pairs: unknown sort");
        set_device(c);
        sgipc::sort_pairs_dispatch(*c, d_keys_in, d_keys_out, d_vals_in, d_vals_out, n, key_bytes, bits, impl);
        if (impl == SGIPC_SORT_OWN && sgipc::radix_sort_error(*c))
            throw std::runtime_error("sort_pairs: radix sort look-back failed");
    });
}

int sgipc_get_batch_mode(sgipc_ctx* c, int* persistent, int* threads) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (persistent) *persistent = c->batch_persistent;
        if (threads) *threads = c->batch_threads;
    });
}

int sgipc_solve(sgipc_ctx* c, const double* b, double* x, const sgipc_solve_opts* o, sgipc_solve_stats* st) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!b || !x) throw sgipc::arg_error("solve: null vector");
        const sgipc::SolveOpts so = solve_opts_of(c, o);
        set_device(c);
        sgipc::SolveStats s2;
        sgipc::prof_begin(*c, sgipc::PROF_SOLVE);
        sgipc::pcg_solve(*c, b, x, so, s2);
        sgipc::prof_end(*c, sgipc::PROF_SOLVE);
        sgipc::prof_flush(*c);
        if (st) {
            st->iterations = s2.iterations;
            st->converged = s2.converged;
            st->refinements = s2.refinements;
            st->rel_residual = s2.rel_residual;
            st->rz0 = s2.rz0;
        }
    });
}

int sgipc_kinetic_energy(sgipc_ctx* c, int n, const double* x, const double* xt, const double* m, double* e) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (n < 0 || !e || (n > 0 && (!x || !xt || !m))) throw sgipc::arg_error("kinetic_energy: invalid argument");
        set_device(c);
        *e = sgipc::kinetic_energy(*c, n, x, xt, m);
    });
}

int sgipc_fem_energy(sgipc_ctx* c, int T, const double* x, const uint32_t* tets, const double* DmInv,
                     const double* vol, const double* mu, const double* lam, double dt, double* e) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (T < 0 || !e || (T > 0 && (!x || !tets || !DmInv || !vol || !mu || !lam)))
            throw sgipc::arg_error("fem_energy: invalid argument");
        set_device(c);
        *e = sgipc::fem_energy(*c, T, x, tets, DmInv, vol, mu, lam, dt);
    });
}

int sgipc_fem_energy_noalloc(sgipc_ctx* c, int T, const double* x, const uint32_t* tets, const double* DmInv,
                             const double* vol, const double* mu, const double* lam, double dt, double* d_e) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (T < 0 || !d_e || (T > 0 && (!x || !tets || !DmInv || !vol || !mu || !lam)))
            throw sgipc::arg_error("fem_energy_noalloc: invalid argument");
        set_device(c);
        sgipc::fem_energy_noalloc(*c, T, x, tets, DmInv, vol, mu, lam, dt, d_e);
    });
}

int sgipc_barrier_energy(sgipc_ctx* c, int np, const int* pairs, const double* x, const double* xr, double dhat,
                         double kappa, double* e) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (np < 0 || !e || (np > 0 && (!pairs || !x || !xr)) || !(dhat > 0.0))
            throw sgipc::arg_error("barrier_energy: invalid argument");
        set_device(c);
        *e = sgipc::barrier_energy(*c, np, pairs, x, xr, dhat, kappa);
    });
}

int sgipc_profile_enable(sgipc_ctx* c, int on) {
    CHECK_CTX(c);
    return guard(c, [&] {
        set_device(c);
        sgipc::prof_flush(*c);
        c->prof.on = on != 0;
    });
}

int sgipc_profile_read(sgipc_ctx* c, int slot, long long* calls, double* total_ms, int reset) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (slot < 0 || slot >= sgipc::PROF_NSLOTS) throw sgipc::arg_error("profile_read: bad slot");
        set_device(c);
        sgipc::prof_flush(*c);
        if (calls) *calls = c->prof.calls[slot];
        if (total_ms) *total_ms = c->prof.ms[slot];
        if (reset) { c->prof.calls[slot] = 0; c->prof.ms[slot] = 0.0; }
    });
}

int sgipc_mas_info(sgipc_ctx* c, int* L, int* total, int* ls) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!c->mas.ready) throw sgipc::state_error("mas_info: MAS not built");
        if (L) *L = c->mas.L;
        if (total) *total = c->mas.total;
        if (ls)
            for (int i = 0; i < 2 * (c->mas.L + 1); ++i) ls[i] = c->mas.level_size_h[i];
    });
}

int sgipc_mas_tables(sgipc_ctx* c, int* cs, int* gn, int* ct, uint32_t* fm, int* p0) {
    CHECK_CTX(c);
    return guard(c, [&] {
        sgipc::MasState& m = c->mas;
        if (!m.ready) throw sgipc::state_error("mas_tables: MAS not built");
        set_device(c);
        SGIPC_HIP(hipStreamSynchronize(c->stream));
        if (cs) SGIPC_HIP(hipMemcpy(cs, m.coarse_space.p, sizeof(int) * (size_t)m.L * m.n, hipMemcpyDeviceToHost));
        if (gn) SGIPC_HIP(hipMemcpy(gn, m.going_next.p, sizeof(int) * (size_t)std::max(m.total, m.n), hipMemcpyDeviceToHost));
        if (ct) SGIPC_HIP(hipMemcpy(ct, m.coarse_table.p, sizeof(int) * 6 * (size_t)m.n, hipMemcpyDeviceToHost));
        if (fm) SGIPC_HIP(hipMemcpy(fm, m.fine_mask.p, sizeof(uint32_t) * (size_t)m.n, hipMemcpyDeviceToHost));
        if (p0) SGIPC_HIP(hipMemcpy(p0, m.prefix0.p, sizeof(int) * (size_t)(m.M / m.bank), hipMemcpyDeviceToHost));
    });
}

int sgipc_mas_bank_inverse(sgipc_ctx* c, int b, double* inv) {
    CHECK_CTX(c);
    return guard(c, [&] {
        sgipc::MasState& m = c->mas;
        if (!m.ready) throw sgipc::state_error("mas_bank_inverse: MAS not built");
        if (!m.inv64_ok) throw sgipc::state_error("mas_bank_inverse: only FP32 inverses were built");
        if (b < 0 || b >= m.nbanks || !inv) throw sgipc::arg_error("mas_bank_inverse: bad bank");
        set_device(c);
        SGIPC_HIP(hipStreamSynchronize(c->stream));
        // device copy: upper triangle packed by rows; expand to dense row-major
        const size_t D = 3 * (size_t)m.bank, NP = D * (D + 1) / 2;
        std::vector<double> up(NP);
        SGIPC_HIP(hipMemcpy(up.data(), m.inv64.p + (size_t)b * NP, sizeof(double) * NP, hipMemcpyDeviceToHost));
        size_t e = 0;
        for (size_t r = 0; r < D; ++r)
            for (size_t q = r; q < D; ++q, ++e) inv[r * D + q] = inv[q * D + r] = up[e];
    });
}

int sgipc_mas_bank_inverse_f32(sgipc_ctx* c, int b, float* inv) {
    CHECK_CTX(c);
    return guard(c, [&] {
        sgipc::MasState& m = c->mas;
        if (!m.ready) throw sgipc::state_error("mas_bank_inverse_f32: MAS not built");
        if (b < 0 || b >= m.nbanks || !inv) throw sgipc::arg_error("mas_bank_inverse_f32: bad bank");
        set_device(c);
        SGIPC_HIP(hipStreamSynchronize(c->stream));
        const size_t D = 3 * (size_t)m.bank, NP = D * (D + 1) / 2;
        std::vector<float> up(NP);
        SGIPC_HIP(hipMemcpy(up.data(), m.inv32.p + (size_t)b * NP, sizeof(float) * NP, hipMemcpyDeviceToHost));
        size_t e = 0;
        for (size_t r = 0; r < D; ++r)
            for (size_t q = r; q < D; ++q, ++e) inv[r * D + q] = inv[q * D + r] = up[e];
    });
}

int sgipc_mas_bank_inverse_f16(sgipc_ctx* c, int b, float* inv) {
    CHECK_CTX(c);
    return guard(c, [&] {
        sgipc::MasState& m = c->mas;
        if (!m.ready) throw sgipc::state_error("mas_bank_inverse_f16: MAS not built");
        if (!m.inv16_ok) throw sgipc::state_error("mas_bank_inverse_f16: no half-storage inverses were built");
        if (b < 0 || b >= m.nbanks || !inv) throw sgipc::arg_error("mas_bank_inverse_f16: bad bank");
        set_device(c);
        SGIPC_HIP(hipStreamSynchronize(c->stream));
        const size_t D = 3 * (size_t)m.bank, NP = D * (D + 1) / 2;
        std::vector<unsigned short> up(NP);
        std::vector<float> d(D);
        SGIPC_HIP(hipMemcpy(up.data(), m.inv16.p + (size_t)b * NP, sizeof(unsigned short) * NP, hipMemcpyDeviceToHost));
        SGIPC_HIP(hipMemcpy(d.data(), m.inv16_d.p + (size_t)b * D, sizeof(float) * D, hipMemcpyDeviceToHost));
        size_t e = 0;
        for (size_t r = 0; r < D; ++r)
            for (size_t q = r; q < D; ++q, ++e) {
                _Float16 h;
                std::memcpy(&h, &up[e], sizeof h);
                inv[r * D + q] = inv[q * D + r] = d[r] * (float)h * d[q];
            }
    });
}

int sgipc_set_bank_storage(sgipc_ctx* c, int store) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (store != SGIPC_BANK_STORE_F32 && store != SGIPC_BANK_STORE_F16)
            throw sgipc::arg_error("set_bank_storage: unknown storage");
        c->bank_store = store;
    });
}

int sgipc_set_bank_inverse(sgipc_ctx* c, int mode) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (mode < SGIPC_BANK_INVERSE_VALU64 || mode > SGIPC_BANK_INVERSE_MFMA32_REFINED)
            throw sgipc::arg_error("set_bank_inverse: unknown mode");
        c->bank_inverse = mode;
    });
}

int sgipc_malloc(sgipc_ctx* c, void** p, size_t bytes) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (!p) throw sgipc::arg_error("malloc: null output");
        set_device(c);
        SGIPC_HIP(hipMalloc(p, bytes ? bytes : 1));
    });
}
int sgipc_free(sgipc_ctx* c, void* p) {
    CHECK_CTX(c);
    return guard(c, [&] { set_device(c); if (p) SGIPC_HIP(hipFree(p)); });
}
int sgipc_memcpy_h2d(sgipc_ctx* c, void* d, const void* h, size_t bytes) {
    CHECK_CTX(c);
    return guard(c, [&] {
        set_device(c);
        SGIPC_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
        SGIPC_HIP(hipStreamSynchronize(c->stream));
    });
}
int sgipc_memcpy_d2h(sgipc_ctx* c, void* h, const void* d, size_t bytes) {
    CHECK_CTX(c);
    return guard(c, [&] {
        set_device(c);
        SGIPC_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
        SGIPC_HIP(hipStreamSynchronize(c->stream));
    });
}
int sgipc_memcpy_d2d(sgipc_ctx* c, void* dst, const void* src, size_t bytes) {
    CHECK_CTX(c);
    return guard(c, [&] {
        if (bytes && (!dst || !src)) throw sgipc::arg_error("memcpy_d2d: null pointer");
        set_device(c);
        if (bytes) SGIPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    });
}

}  // extern "C"
