// gipc_linear_system.hpp -- Stiff-GIPC's linear-system API (namespace gipc,
// StiffGIPC/linear_system/) over the sgipc C ABI (sgipc.h), header-only.
//
// A GIPC build that includes this header instead of the reference's
// linear_system headers keeps its calls: GlobalLinearSystem::create<T>()
// registers subsystems, one solver and the preconditioners, the caller
// assembles triplets into *gipc_global_triplet and advances its
// global_triplet_offset, and solve_linear_system() builds the system (RHS
// from the subsystems, in-place triplet -> BCOO conversion, preconditioners),
// solves it and hands each subsystem its part of the solution.  Everything
// that runs on the device runs in libsgipc (HIP kernels for gfx950).
//
// Mapping to the reference (paths under StiffGIPC/linear_system/):
//   GlobalLinearSystem        linear_system/global_linear_system.h:14-104, .cu:8-197
//   ILinearSubsystem,
//   DiagonalSubsystem         linear_system/linear_subsystem.h:14-149, .cu:23-29
//   IterativeSolver           linear_system/i_linear_system_solver.h:9-59
//   IPreconditioner,
//   LocalPreconditioner,
//   GlobalPreconditioner      linear_system/i_preconditioner.h:18-98
//   GIPCTripletMatrix         linear_system/global_matrix.h:18-203 (the caller's
//                             device arrays; values 9 doubles column-major)
//   PCGSolverConfig/PCGSolver solver/pcg_solver.h:6-16, pcg_solver.cu:158-267
//   MAS_Preconditioner        preconditioner/fem_mas_preconditioner.{h,cu}
//   ABDPreconditioner         preconditioner/abd_preconditioner.{h,cu}
//   DiagPreconditioner        preconditioner/diag_preconditioner.{h,cu}
//   ::MASPreconditioner       MASPreconditioner.cuh:57-102 (here: the host-side
//                             partition inputs; the MAS itself lives in the library)
//
// Differences a caller sees:
//   - errors throw gipc::Error (the reference aborts);
//   - vector views are gipc::DenseVectorView / CDenseVectorView (device
//     pointer + offset + size) instead of muda's;
//   - ctx() of subsystems / solvers / preconditioners is the library context
//     (sgipc_ctx*), which the sgipc_* entry points take;
//   - the library preconditioners (MAS, ABD block-Jacobi, Diag) are realised
//     by one preconditioner build and one apply inside the library, and
//     PCGSolver runs the library's device-resident PCG with them; a
//     preconditioner a caller derives itself is applied by
//     GlobalLinearSystem's apply_preconditioner() for the caller's own
//     IterativeSolver, and a system may not mix the two kinds;
//   - the reference's PCGSolverConfig::use_bsr has no effect (the library
//     has one SpMV format), and the config has the library's options too
//     (stop rule, precision, iteration cap).
#ifndef GIPC_LINEAR_SYSTEM_HPP
#define GIPC_LINEAR_SYSTEM_HPP

#include <cstddef>
#include <cstdint>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "sgipc.h"

namespace gipc {

using Float = double;
using SizeT = std::size_t;
using IndexT = int;

struct Vector2i {
    IndexT v[2] = {0, 0};
    IndexT operator[](int i) const { return v[i]; }
    IndexT operator()(int i) const { return v[i]; }
};

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// ctx: the context of the failed call, or NULL for context-free calls
// (their message is the thread's last context-free error)
inline void check(int rc, const sgipc_ctx* ctx, const char* where) {
    if (rc == SGIPC_OK) return;
    std::string msg = where;
    msg += ": ";
    msg += sgipc_last_error(ctx);
    throw Error(rc, msg);
}

// Device vector view: a device pointer, the view's offset in the vector it
// was taken from, and its size (muda::DenseVectorView<Float> in the reference).
template <typename T>
class DenseVectorViewT {
public:
    DenseVectorViewT() = default;
    DenseVectorViewT(T* data, SizeT size, SizeT offset = 0) : data_(data), size_(size), offset_(offset) {}
    template <typename U, typename = std::enable_if_t<std::is_same_v<T, const U>>>
    DenseVectorViewT(const DenseVectorViewT<U>& o) : data_(o.data()), size_(o.size()), offset_(o.offset()) {}
    T* data() const { return data_; }
    SizeT size() const { return size_; }
    // start of this view in the global vector
    SizeT offset() const { return offset_; }
    DenseVectorViewT subview(SizeT off, SizeT n) const {
        if (off + n > size_) throw Error(SGIPC_ERR_INVALID_ARG, "subview out of range");
        return DenseVectorViewT(data_ + off, n, offset_ + off);
    }

private:
    T* data_ = nullptr;
    SizeT size_ = 0;
    SizeT offset_ = 0;
};
using DenseVectorView = DenseVectorViewT<Float>;
using CDenseVectorView = DenseVectorViewT<const Float>;

// The caller's triplet store: device arrays the assembly writes into (9
// doubles per 3x3 block, column-major; int32 block rows / cols), converted
// in place to the unique upper BCOO at [0, h_unique_key_number).
class GIPCTripletMatrix {
public:
    GIPCTripletMatrix() = default;
    GIPCTripletMatrix(Float* d_values, int* d_rows, int* d_cols, long long capacity) {
        bind(d_values, d_rows, d_cols, capacity);
    }
    void bind(Float* d_values, int* d_rows, int* d_cols, long long capacity) {
        values_ = d_values;
        rows_ = d_rows;
        cols_ = d_cols;
        capacity_ = capacity;
    }
    Float* block_values() const { return values_; }
    int* block_row_indices() const { return rows_; }
    int* block_col_indices() const { return cols_; }
    long long capacity() const { return capacity_; }

    long long global_triplet_offset = 0;  // triplets assembled so far
    int h_unique_key_number = 0;          // unique blocks after the conversion

private:
    Float* values_ = nullptr;
    int* rows_ = nullptr;
    int* cols_ = nullptr;
    long long capacity_ = 0;
};

class GlobalLinearSystem;

class ILinearSubsystem {
public:
    using DenseVectorView = gipc::DenseVectorView;
    using CDenseVectorView = gipc::CDenseVectorView;

    ILinearSubsystem() = default;
    virtual ~ILinearSubsystem() = default;
    ILinearSubsystem(const ILinearSubsystem&) = delete;
    ILinearSubsystem& operator=(const ILinearSubsystem&) = delete;

protected:
    // report the subsystem's size (DiagonalSubsystem: right_hand_side_dof)
    virtual void report_subsystem_info() = 0;
    // the library context the sgipc_* calls of a subsystem take
    sgipc_ctx* ctx() const;

private:
    friend class GlobalLinearSystem;
    friend class DiagonalSubsystem;
    friend class LocalPreconditioner;
    GlobalLinearSystem* m_system = nullptr;
    IndexT m_hid = 0, m_gid = 0;
    virtual Vector2i dof_offset() const = 0;
    virtual void do_assemble(DenseVectorView gradient) = 0;
};

// A subrange of the solution / right-hand side (and a diagonal block of the
// Hessian) of the global system.
class DiagonalSubsystem : public ILinearSubsystem {
public:
    DiagonalSubsystem() = default;
    virtual ~DiagonalSubsystem() = default;

protected:
    // size of this subsystem's right-hand side (a multiple of 3)
    void right_hand_side_dof(SizeT dof) {
        if (dof % 3 != 0) throw Error(SGIPC_ERR_INVALID_ARG, "right_hand_side_dof must be a multiple of 3");
        m_right_hand_side_dof = dof;
    }
    SizeT right_hand_side_dof() const { return m_right_hand_side_dof; }
    // fill gradient[0, right_hand_side_dof()) (gradient.offset() is the
    // global start)
    virtual void assemble(DenseVectorView gradient) = 0;
    // read this subsystem's part of the solution
    virtual void retrieve_solution(CDenseVectorView dx) = 0;

private:
    friend class GlobalLinearSystem;
    friend class LocalPreconditioner;
    SizeT m_right_hand_side_dof = 0;
    IndexT m_dof_offset = 0;
    Vector2i dof_offset() const final { return Vector2i{{m_dof_offset, m_dof_offset}}; }
    void do_assemble(DenseVectorView gradient) final {
        assemble(gradient.subview((SizeT)m_dof_offset, m_right_hand_side_dof));
    }
    void do_retrieve_solution(CDenseVectorView dx) {
        retrieve_solution(dx.subview((SizeT)m_dof_offset, m_right_hand_side_dof));
    }
};

class IterativeSolver {
public:
    IterativeSolver() = default;
    virtual ~IterativeSolver() = default;
    IterativeSolver(const IterativeSolver&) = delete;
    IterativeSolver& operator=(const IterativeSolver&) = delete;

protected:
    // solve A x = b; returns the iteration count (0 for a direct solver)
    virtual SizeT solve(DenseVectorView x, CDenseVectorView b) = 0;
    // y = a A x + b y on the converted system
    void spmv(Float a, CDenseVectorView x, Float b, DenseVectorView y);
    void spmv(CDenseVectorView x, DenseVectorView y) { spmv(1.0, x, 0.0, y); }
    // z = M^-1 r with the system's preconditioners
    void apply_preconditioner(DenseVectorView z, CDenseVectorView r) const;
    sgipc_ctx* ctx() const;
    GlobalLinearSystem& system() const { return *m_system; }

private:
    friend class GlobalLinearSystem;
    GlobalLinearSystem* m_system = nullptr;
};

class IPreconditioner {
public:
    IPreconditioner() = default;
    virtual ~IPreconditioner() = default;
    IPreconditioner(const IPreconditioner&) = delete;
    IPreconditioner& operator=(const IPreconditioner&) = delete;

protected:
    sgipc_ctx* ctx() const;
    GlobalLinearSystem& system() const { return *m_system; }
    // library preconditioners are built and applied by the library as one
    // preconditioner (see the header comment)
    virtual bool in_library() const { return false; }

private:
    friend class GlobalLinearSystem;
    friend class LocalPreconditioner;
    friend class GlobalPreconditioner;
    GlobalLinearSystem* m_system = nullptr;
    virtual void do_apply(CDenseVectorView r, DenseVectorView z) = 0;
    virtual void do_assemble(GIPCTripletMatrix& global_triplets) = 0;
};

// Preconditioner of one subsystem's range.  preconditioner_id 0 assembles
// before the conversion (the reference's ABD preconditioner), others after.
class LocalPreconditioner : public IPreconditioner {
public:
    explicit LocalPreconditioner(DiagonalSubsystem& subsystem) : m_subsystem(&subsystem) {}
    virtual ~LocalPreconditioner() = default;
    // first block row of the subsystem
    int get_offset() const { return m_subsystem->m_dof_offset / 3; }
    DiagonalSubsystem& subsystem() const { return *m_subsystem; }
    int preconditioner_id = -1;

protected:
    virtual void assemble() {}
    virtual void apply(CDenseVectorView r, DenseVectorView z) = 0;
    // the subsystem's right-hand side size
    SizeT subsystem_dof() const { return m_subsystem->m_right_hand_side_dof; }

private:
    friend class GlobalLinearSystem;
    DiagonalSubsystem* m_subsystem;
    void do_apply(CDenseVectorView r, DenseVectorView z) override {
        const SizeT off = (SizeT)m_subsystem->m_dof_offset, n = m_subsystem->m_right_hand_side_dof;
        apply(r.subview(off, n), z.subview(off, n));
    }
    void do_assemble(GIPCTripletMatrix&) override { assemble(); }
};

// Preconditioner of the whole system.
class GlobalPreconditioner : public IPreconditioner {
public:
    GlobalPreconditioner() = default;
    virtual ~GlobalPreconditioner() = default;

protected:
    virtual void assemble(GIPCTripletMatrix& global_triplets) = 0;
    virtual void apply(CDenseVectorView r, DenseVectorView z) = 0;

private:
    friend class GlobalLinearSystem;
    void do_apply(CDenseVectorView r, DenseVectorView z) override { apply(r, z); }
    void do_assemble(GIPCTripletMatrix& t) override { assemble(t); }
};

class PCGSolverConfig {
public:
    // maximum iterations = dof * max_iter_ratio
    Float max_iter_ratio = 0.3;
    // reference stop rule: |r.z| <= global_tol_rate * r0.z0
    Float global_tol_rate = 1e-4;
    bool use_bsr = true;  // kept for source compatibility; no effect
    // library options
    int stop_rule = SGIPC_STOP_REF_RZ;      // or SGIPC_STOP_TRUE_REL (||r|| <= tol ||b||)
    int precision = SGIPC_PRECISION_FP64;   // or SGIPC_PRECISION_MIXED
    int max_iter = 0;                       // > 0 overrides the ratio
    int check_every = 8;                    // iterations between host polls
};

class PCGSolver : public IterativeSolver {
public:
    explicit PCGSolver(const PCGSolverConfig& cfg) : m_config(cfg) {}
    virtual ~PCGSolver() = default;
    void config(const PCGSolverConfig& cfg) { m_config = cfg; }
    const PCGSolverConfig& config() const { return m_config; }
    const sgipc_solve_stats& last_stats() const { return m_stats; }

protected:
    SizeT solve(DenseVectorView x, CDenseVectorView b) override;

private:
    PCGSolverConfig m_config;
    sgipc_solve_stats m_stats{};
};

}  // namespace gipc

// The reference's MASPreconditioner (a global-namespace class) holds the MAS
// partition and a pointer to the contact pairs it reads its connectivity
// from.  Here it holds the host-side inputs of the library's MAS: vertex
// neighbours and bank slot maps (from the FEM tets and ascending part ids),
// the bank size, the vertex offset of the FEM vertices in the contact pairs'
// vertex numbering, and the device pointer of the pairs (int4, reference
// encoding; their count comes from MAS_Preconditioner's cpNum).
class MASPreconditioner {
public:
    // initPreconditioner_Neighbor (MASPreconditioner.cu:2264-2302) with the
    // host preprocessing of gl_main.cu:1066-1093 (setMAS_partition) and
    // load_mesh.cpp:934-1040 (getVertNeighbors)
    void init_partition(int n_fem, int n_tets, const uint32_t* h_tets, const int* h_part_id, int bank = 16,
                        int collision_node_offset = 0, const int* d_collision_pairs = nullptr) {
        std::vector<uint32_t> start(n_fem > 0 ? n_fem : 1), num(n_fem > 0 ? n_fem : 1),
            list(12 * (size_t)n_tets + 1);
        int len = 0;
        gipc::check(sgipc_build_vertex_neighbors(n_fem, n_tets, h_tets, start.data(), num.data(), list.data(),
                                                 (int)list.size(), &len),
                    nullptr, "sgipc_build_vertex_neighbors");
        const int cap = n_fem > 0 ? bank * (h_part_id[n_fem - 1] + 1) : 0;
        std::vector<int> part_map(cap > 0 ? cap : 1), real_map(n_fem > 0 ? n_fem : 1);
        int M = 0;
        gipc::check(sgipc_build_bank_map(n_fem, h_part_id, bank, part_map.data(), cap, &M, real_map.data()), nullptr,
                    "sgipc_build_bank_map");
        part_map.resize(M > 0 ? M : 1);
        list.resize(len > 0 ? len : 1);
        m_n = n_fem;
        m_bank = bank;
        m_node_offset = collision_node_offset;
        m_start = std::move(start);
        m_num = std::move(num);
        m_list = std::move(list);
        m_list_len = len;
        m_part_map = std::move(part_map);
        m_M = M;
        m_real_map = std::move(real_map);
        m_pairs = d_collision_pairs;
        ++m_generation;
    }
    void collision_pairs(const int* d_pairs) { m_pairs = d_pairs; }
    const int* collision_pairs() const { return m_pairs; }
    int bank() const { return m_bank; }
    int vertex_count() const { return m_n; }

    // upload the topology to a library context (once per context and partition)
    void upload(sgipc_ctx* ctx) {
        if (m_uploaded_ctx == ctx && m_uploaded_generation == m_generation) return;
        if (m_generation == 0) throw gipc::Error(SGIPC_ERR_STATE, "MASPreconditioner: init_partition first");
        gipc::check(sgipc_set_mas_topology(ctx, m_n, m_node_offset, m_start.data(), m_num.data(), m_list.data(),
                                           m_list_len, m_part_map.data(), m_M, m_real_map.data(), m_bank),
                    ctx, "sgipc_set_mas_topology");
        m_uploaded_ctx = ctx;
        m_uploaded_generation = m_generation;
    }

private:
    int m_n = 0, m_bank = 16, m_node_offset = 0, m_list_len = 0, m_M = 0;
    std::vector<uint32_t> m_start, m_num, m_list;
    std::vector<int> m_part_map, m_real_map;
    const int* m_pairs = nullptr;
    long long m_generation = 0, m_uploaded_generation = -1;
    sgipc_ctx* m_uploaded_ctx = nullptr;
};

namespace gipc {

class GlobalLinearSystem {
    template <typename T>
    using U = std::unique_ptr<T>;

public:
    static constexpr int BlockSize = 3;

    // device: HIP device index; hip_stream: a caller stream for all work
    // (default: the library context's own stream)
    explicit GlobalLinearSystem(int device = 0, void* hip_stream = nullptr) {
        check(sgipc_ctx_create(device, &m_ctx), nullptr, "sgipc_ctx_create");
        if (hip_stream) check(sgipc_ctx_set_stream(m_ctx, hip_stream), m_ctx, "sgipc_ctx_set_stream");
    }
    ~GlobalLinearSystem() {
        if (m_ctx) {
            if (m_x) sgipc_free(m_ctx, m_x);
            if (m_b) sgipc_free(m_ctx, m_b);
            sgipc_ctx_destroy(m_ctx);
        }
    }
    GlobalLinearSystem(const GlobalLinearSystem&) = delete;
    GlobalLinearSystem& operator=(const GlobalLinearSystem&) = delete;

    // register a subsystem (DiagonalSubsystem), the solver (IterativeSolver)
    // or a preconditioner (Local / Global); returns the created object
    template <typename T, typename... Args>
    T& create(Args&&... args) {
        if constexpr (std::is_base_of_v<DiagonalSubsystem, T>) {
            auto p = std::make_unique<T>(std::forward<Args>(args)...);
            T& ref = *p;
            ILinearSubsystem& base = ref;
            base.m_system = this;
            base.m_gid = (IndexT)m_subsystems.size();
            base.m_hid = (IndexT)m_subsystems.size();
            m_subsystems.push_back(std::move(p));
            return ref;
        } else if constexpr (std::is_base_of_v<IterativeSolver, T>) {
            auto p = std::make_unique<T>(std::forward<Args>(args)...);
            T& ref = *p;
            static_cast<IterativeSolver&>(ref).m_system = this;
            m_solver = std::move(p);
            return ref;
        } else if constexpr (std::is_base_of_v<LocalPreconditioner, T>) {
            auto p = std::make_unique<T>(std::forward<Args>(args)...);
            T& ref = *p;
            static_cast<IPreconditioner&>(ref).m_system = this;
            m_local_preconditioners.push_back(std::move(p));
            return ref;
        } else if constexpr (std::is_base_of_v<GlobalPreconditioner, T>) {
            if (m_global_preconditioner) throw Error(SGIPC_ERR_STATE, "a global preconditioner already exists");
            auto p = std::make_unique<T>(std::forward<Args>(args)...);
            T& ref = *p;
            static_cast<IPreconditioner&>(ref).m_system = this;
            m_global_preconditioner = std::move(p);
            return ref;
        } else {
            static_assert(std::is_base_of_v<DiagonalSubsystem, T>, "create<T>: unknown type");
        }
    }

    // Build the system from the subsystems (RHS, conversion of the triplets
    // assembled into *gipc_global_triplet, preconditioners), solve it with
    // the registered solver and distribute the solution; returns the
    // iteration count, 0 when the system is empty.
    SizeT solve_linear_system() {
        if (!build_linear_system()) return 0;
        if (!m_solver) throw Error(SGIPC_ERR_STATE, "solve_linear_system: no solver (create<PCGSolver>)");
        SizeT it = m_solver->solve(DenseVectorView(m_x, m_n), CDenseVectorView(m_b, m_n));
        distribute_solution();
        return it;
    }

    // Independent scenes assembled at disjoint block-row ranges of one
    // subsystem: scene s owns block rows [row_start[s], row_start[s+1]) and
    // runs its own PCG in shared launches (library extension, no reference
    // counterpart).  Per-scene statistics in `stats`.
    void solve_linear_system_batched(const std::vector<int>& row_start, std::vector<sgipc_solve_stats>& stats) {
        const int S = (int)row_start.size() - 1;
        if (S <= 0) throw Error(SGIPC_ERR_INVALID_ARG, "solve_linear_system_batched: no scenes");
        stats.assign(S, sgipc_solve_stats{});
        if (!build_linear_system()) return;
        auto* pcg = dynamic_cast<PCGSolver*>(m_solver.get());
        PCGSolverConfig cfg = pcg ? pcg->config() : PCGSolverConfig{};
        sgipc_solve_opts o{cfg.precision, cfg.stop_rule, cfg.global_tol_rate, cfg.max_iter_ratio, cfg.max_iter,
                           cfg.check_every};
        check(sgipc_solve_batched(m_ctx, S, row_start.data(), m_b, m_x, &o, stats.data()), m_ctx,
              "sgipc_solve_batched");
        distribute_solution();
    }

    // the caller's triplet store (set by the caller, as GIPC does)
    GIPCTripletMatrix* gipc_global_triplet = nullptr;

    // library context (for the sgipc_* assembly entry points)
    sgipc_ctx* handle() const { return m_ctx; }
    // register *gipc_global_triplet's arrays with the library (done by
    // solve_linear_system and by gipc::Assembler before they write)
    void bind_triplets() {
        if (!gipc_global_triplet) throw Error(SGIPC_ERR_STATE, "gipc_global_triplet is not set");
        GIPCTripletMatrix& t = *gipc_global_triplet;
        if (t.block_values() == m_bound_vals && t.block_row_indices() == m_bound_rows &&
            t.block_col_indices() == m_bound_cols && t.capacity() == m_bound_cap)
            return;
        check(sgipc_triplets_view(m_ctx, t.block_values(), t.block_row_indices(), t.block_col_indices(), t.capacity()),
              m_ctx, "sgipc_triplets_view");
        m_bound_vals = t.block_values();
        m_bound_rows = t.block_row_indices();
        m_bound_cols = t.block_col_indices();
        m_bound_cap = t.capacity();
    }
    SizeT dof() const { return m_n; }
    // z = M^-1 r with the registered preconditioners (the solver helper)
    void apply_preconditioner(DenseVectorView z, CDenseVectorView r) {
        if (preconditioners_in_library()) {
            check(sgipc_precond_apply(m_ctx, r.data(), z.data()), m_ctx, "sgipc_precond_apply");
            return;
        }
        if (m_global_preconditioner) m_global_preconditioner->do_apply(r, z);
        else check(sgipc_memcpy_d2d(m_ctx, z.data(), r.data(), sizeof(Float) * r.size()), m_ctx, "sgipc_memcpy_d2d");
        for (auto& p : m_local_preconditioners) p->do_apply(r, z);
    }
    void spmv(Float a, CDenseVectorView x, Float b, DenseVectorView y) {
        check(sgipc_spmv(m_ctx, a, x.data(), b, y.data()), m_ctx, "sgipc_spmv");
    }
    // true when every registered preconditioner is a library one (or none)
    bool preconditioners_in_library() const {
        if (m_global_preconditioner && !in_library(*m_global_preconditioner)) return false;
        for (auto& p : m_local_preconditioners)
            if (!in_library(*p)) return false;
        return true;
    }

private:
    sgipc_ctx* m_ctx = nullptr;
    std::vector<U<DiagonalSubsystem>> m_subsystems;
    std::vector<U<LocalPreconditioner>> m_local_preconditioners;
    U<GlobalPreconditioner> m_global_preconditioner;
    U<IterativeSolver> m_solver;
    Float* m_x = nullptr;
    Float* m_b = nullptr;
    SizeT m_n = 0, m_cap = 0;
    const Float* m_bound_vals = nullptr;
    const int* m_bound_rows = nullptr;
    const int* m_bound_cols = nullptr;
    long long m_bound_cap = -1;

    static bool in_library(const IPreconditioner& p) { return p.in_library(); }

    void reserve_vectors(SizeT n) {
        if (n <= m_cap) return;
        if (m_x) { sgipc_free(m_ctx, m_x); m_x = nullptr; }
        if (m_b) { sgipc_free(m_ctx, m_b); m_b = nullptr; }
        void* px = nullptr;
        void* pb = nullptr;
        check(sgipc_malloc(m_ctx, &px, sizeof(Float) * n), m_ctx, "sgipc_malloc");
        m_x = static_cast<Float*>(px);
        check(sgipc_malloc(m_ctx, &pb, sizeof(Float) * n), m_ctx, "sgipc_malloc");
        m_b = static_cast<Float*>(pb);
        m_cap = n;
    }

    // global_linear_system.cu:8-74
    bool build_linear_system() {
        if (!gipc_global_triplet) throw Error(SGIPC_ERR_STATE, "gipc_global_triplet is not set");
        for (auto& s : m_subsystems) s->report_subsystem_info();
        SizeT total = 0;
        for (auto& s : m_subsystems) {
            s->m_dof_offset = (IndexT)total;
            total += s->m_right_hand_side_dof;
        }
        m_n = total;
        if (gipc_global_triplet->global_triplet_offset == 0 || total == 0) return false;
        reserve_vectors(total);
        const DenseVectorView b(m_b, total);
        for (auto& s : m_subsystems) s->do_assemble(b);
        bool any_library = m_global_preconditioner && in_library(*m_global_preconditioner);
        for (auto& p : m_local_preconditioners) any_library |= in_library(*p);
        if (any_library && !preconditioners_in_library())
            throw Error(SGIPC_ERR_STATE,
                        "library preconditioners cannot be combined with preconditioners defined by the caller");
        size_t start = 0;
        if (!m_local_preconditioners.empty() && m_local_preconditioners[0]->preconditioner_id == 0) {
            m_local_preconditioners[0]->assemble();
            start = 1;
        }
        bind_triplets();
        int K = 0;
        check(sgipc_convert(m_ctx, gipc_global_triplet->global_triplet_offset, (int)(total / 3), &K), m_ctx,
              "sgipc_convert");
        gipc_global_triplet->h_unique_key_number = K;
        if (m_global_preconditioner) m_global_preconditioner->do_assemble(*gipc_global_triplet);
        for (size_t i = start; i < m_local_preconditioners.size(); ++i) m_local_preconditioners[i]->assemble();
        return true;
    }

    // global_linear_system.cu:76-84
    void distribute_solution() {
        const CDenseVectorView x(m_x, m_n);
        for (auto& s : m_subsystems) s->do_retrieve_solution(x);
        check(sgipc_ctx_synchronize(m_ctx), m_ctx, "sgipc_ctx_synchronize");
    }
};

inline sgipc_ctx* ILinearSubsystem::ctx() const { return m_system->handle(); }
inline sgipc_ctx* IterativeSolver::ctx() const { return m_system->handle(); }
inline sgipc_ctx* IPreconditioner::ctx() const { return m_system->handle(); }
inline void IterativeSolver::spmv(Float a, CDenseVectorView x, Float b, DenseVectorView y) {
    m_system->spmv(a, x, b, y);
}
inline void IterativeSolver::apply_preconditioner(DenseVectorView z, CDenseVectorView r) const {
    m_system->apply_preconditioner(z, r);
}

// pcg_solver.cu:158-267: the library's device-resident PCG with the
// preconditioner the library preconditioners built (dof * max_iter_ratio
// iterations at most).
inline SizeT PCGSolver::solve(DenseVectorView x, CDenseVectorView b) {
    if (!system().preconditioners_in_library())
        throw Error(SGIPC_ERR_STATE,
                    "PCGSolver runs inside the library with its preconditioners; a preconditioner defined by the "
                    "caller needs an IterativeSolver defined by the caller");
    sgipc_solve_opts o{m_config.precision, m_config.stop_rule, m_config.global_tol_rate, m_config.max_iter_ratio,
                       m_config.max_iter, m_config.check_every};
    check(sgipc_solve(ctx(), b.data(), x.data(), &o, &m_stats), ctx(), "sgipc_solve");
    return (SizeT)m_stats.iterations;
}

// MAS on an FEM subsystem's rows (fem_mas_preconditioner.cu:6-37):
// (subsystem, the partition holder, masses (unused, as in the reference),
// host pointer to the contact-pair count).  preconditioner_id 1.
class MAS_Preconditioner : public LocalPreconditioner {
public:
    MAS_Preconditioner(DiagonalSubsystem& subsystem, ::MASPreconditioner& mas, Float* masses, uint32_t* cpNum,
                       int kind = SGIPC_PRECOND_MAS)
        : LocalPreconditioner(subsystem), m_mas(mas), m_masses(masses), m_cpNum(cpNum), m_kind(kind) {
        preconditioner_id = 1;
        if (kind != SGIPC_PRECOND_MAS && kind != SGIPC_PRECOND_MAS_FP32)
            throw Error(SGIPC_ERR_INVALID_ARG, "MAS_Preconditioner: kind must be SGIPC_PRECOND_MAS(_FP32)");
    }
    void kind(int k) { m_kind = k; }

protected:
    bool in_library() const override { return true; }
    void assemble() override {
        m_mas.upload(ctx());
        const int npairs = m_cpNum ? (int)m_cpNum[0] : 0;
        const int sub_nb = (int)(subsystem_dof() / 3);
        check(sgipc_precond_build(ctx(), m_kind, get_offset(), sub_nb, npairs ? m_mas.collision_pairs() : nullptr,
                                  npairs),
              ctx(), "sgipc_precond_build");
    }
    // applied by the library together with the other library preconditioners
    void apply(CDenseVectorView, DenseVectorView) override {}

private:
    ::MASPreconditioner& m_mas;
    Float* m_masses;
    uint32_t* m_cpNum;
    int m_kind;
};

// 12x12 block-Jacobi of the affine bodies (abd_preconditioner.cu:7-38):
// built by the library with the MAS of the FEM rows after the bodies' rows
// (the bodies are set up with sgipc_abd_setup).  The reference passes its
// ABDSystem and ABDSimData too; the library holds the body data, so any
// further arguments are accepted and unused.  preconditioner_id 0.
class ABDPreconditioner : public LocalPreconditioner {
public:
    template <typename... SimulatorState>
    explicit ABDPreconditioner(DiagonalSubsystem& subsystem, SimulatorState&...) : LocalPreconditioner(subsystem) {
        preconditioner_id = 0;
    }

protected:
    bool in_library() const override { return true; }
    void assemble() override {}
    void apply(CDenseVectorView, DenseVectorView) override {}
};

// 3x3 block-Jacobi over the whole system (diag_preconditioner.cu:9-66).
class DiagPreconditioner : public GlobalPreconditioner {
public:
    DiagPreconditioner() = default;

protected:
    bool in_library() const override { return true; }
    void assemble(GIPCTripletMatrix&) override {
        check(sgipc_precond_build(ctx(), SGIPC_PRECOND_DIAG, 0, (int)(system().dof() / 3), nullptr, 0), ctx(),
              "sgipc_precond_build");
    }
    void apply(CDenseVectorView, DenseVectorView) override {}
};

// Assembly into *gipc_global_triplet through the library, in the order
// GIPC::computeGradientAndHessian (GIPC.cu:10110-10344) writes: each call
// appends at global_triplet_offset and advances it.  (The reference's
// GIPC assembles with its own kernels; this is the library's replacement.)
class Assembler {
public:
    explicit Assembler(GlobalLinearSystem& sys, int n_bodies = 0) : m_sys(sys), m_nb(n_bodies) {}
    GIPCTripletMatrix& triplets() const { return *m_sys.gipc_global_triplet; }
    void begin() {
        m_sys.bind_triplets();
        triplets().global_triplet_offset = 0;
    }
    void kinetic_gradient(int n, const Float* d_x, const Float* d_xtilde, const Float* d_mass, Float* d_shape) {
        check(sgipc_kinetic_gradient(ctx(), n, d_x, d_xtilde, d_mass, d_shape), ctx(), "sgipc_kinetic_gradient");
    }
    void fem(int T, const Float* d_x, const uint32_t* d_tets, const Float* d_DmInv, const Float* d_vol,
             const Float* d_mu, const Float* d_lambda, Float dt, const int* d_btype, Float* d_shape) {
        m_sys.bind_triplets();
        check(sgipc_assemble_fem(ctx(), T, d_x, d_tets, d_DmInv, d_vol, d_mu, d_lambda, dt, d_btype, 4 * m_nb,
                                 triplets().global_triplet_offset, d_shape),
              ctx(), "sgipc_assemble_fem");
        triplets().global_triplet_offset += 10LL * T;
    }
    void mass_diagonal(int n, const Float* d_mass) {
        m_sys.bind_triplets();
        check(sgipc_mass_triplets(ctx(), n, d_mass, 4 * m_nb, triplets().global_triplet_offset), ctx(),
              "sgipc_mass_triplets");
        triplets().global_triplet_offset += n;
    }
    // contact barrier of FEM-only scenes (vertex space = DOF space)
    void barrier(int n_pairs, const int* d_pairs, const int* d_mat_index, const int counts[3], const Float* d_x,
                 const Float* d_xrest, Float dhat, Float kappa, const int* d_btype, Float* d_fb) {
        m_sys.bind_triplets();
        check(sgipc_assemble_barrier(ctx(), n_pairs, d_pairs, d_mat_index, counts, d_x, d_xrest, dhat, kappa, d_btype,
                                     0, triplets().global_triplet_offset, d_fb),
              ctx(), "sgipc_assemble_barrier");
        triplets().global_triplet_offset += 16LL * counts[2] + 9LL * counts[1] + 4LL * counts[0];
    }
    // contact barrier with affine bodies: vertex-space blocks written at the
    // end of the store, mapped to DOF space at the current offset
    void barrier_with_bodies(int n_pairs, const int* d_pairs, const int* d_mat_index, const int counts[3],
                             const Float* d_x, const Float* d_xrest, Float dhat, Float kappa, const int* d_btype,
                             Float* d_fb) {
        m_sys.bind_triplets();
        const long long m = 16LL * counts[2] + 9LL * counts[1] + 4LL * counts[0];
        const long long src = triplets().capacity() - m;
        check(sgipc_assemble_barrier(ctx(), n_pairs, d_pairs, d_mat_index, counts, d_x, d_xrest, dhat, kappa, d_btype,
                                     0, src, d_fb),
              ctx(), "sgipc_assemble_barrier");
        long long n = 0;
        check(sgipc_abd_map_contacts(ctx(), src, m, triplets().global_triplet_offset, &n), ctx(),
              "sgipc_abd_map_contacts");
        triplets().global_triplet_offset += n;
    }
    // ground plane n.x = offset: close vertices of d_cand, then their blocks
    int ground_detect(int n_cand, const uint32_t* d_cand, const Float* d_x, const Float normal[3], Float offset,
                      Float dhat, uint32_t* d_close) {
        int k = 0;
        check(sgipc_ground_detect(ctx(), n_cand, d_cand, d_x, normal, offset, dhat, d_close, &k), ctx(),
              "sgipc_ground_detect");
        return k;
    }
    void ground(int n_close, const uint32_t* d_close, const Float* d_x, const Float normal[3], Float offset,
                Float dhat, Float kappa, const int* d_btype, Float* d_fb) {
        m_sys.bind_triplets();
        check(sgipc_assemble_ground(ctx(), n_close, d_close, d_x, normal, offset, dhat, kappa, d_btype, 0,
                                    triplets().global_triplet_offset, d_fb),
              ctx(), "sgipc_assemble_ground");
        triplets().global_triplet_offset += n_close;
    }
    // affine bodies: body gradient into d_grad_q (12 per body) and body
    // blocks; then J^T fb of the body vertices added to d_grad_q
    void bodies(const Float* d_q, const Float* d_q_tilde, Float kappa_abd, Float dt, const Float* d_fb,
                Float* d_grad_q) {
        m_sys.bind_triplets();
        check(sgipc_abd_assemble(ctx(), d_q, d_q_tilde, kappa_abd, dt, triplets().global_triplet_offset, d_grad_q),
              ctx(), "sgipc_abd_assemble");
        triplets().global_triplet_offset += 10LL * m_nb;
        if (d_fb) check(sgipc_abd_contact_gradient(ctx(), d_fb, d_grad_q), ctx(), "sgipc_abd_contact_gradient");
    }
    void body_positions(const Float* d_q, Float* d_x) {
        check(sgipc_abd_positions(ctx(), d_q, d_x), ctx(), "sgipc_abd_positions");
    }
    // FEMLinearSubsystem::assemble (fem_linear_subsystem.cu:56-84):
    // b = fb + shape gradient, zero on pinned vertices
    void fem_rhs(int n, const Float* d_fb, const Float* d_shape, const int* d_btype, Float* d_b) {
        check(sgipc_fem_rhs(ctx(), n, d_fb, d_shape, d_btype, d_b), ctx(), "sgipc_fem_rhs");
    }

private:
    GlobalLinearSystem& m_sys;
    int m_nb;
    sgipc_ctx* ctx() const { return m_sys.handle(); }
};

}  // namespace gipc

#endif
