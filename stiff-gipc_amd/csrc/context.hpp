// Library context: one HIP stream, caller-owned triplet store, and all
// device workspaces of the Newton-step path (FEM assembly, BCOO, BSR, MAS,
// PCG).  Not thread-safe: one host thread drives a context at a time.
#pragma once
#include "common.hpp"
#include <string>
#include <vector>

namespace sgipc {

// Symmetric SpMV structure over the upper BCOO resident in the triplet view.
// Block row j owns entries ent[eptr[j] .. eptr[j+1]) = {upper block index k,
// column}: first the references to the upper blocks (i, j), i < j, applied
// transposed (column i < j, sorted by k), then its own upper blocks (j, c),
// c >= j, in BCOO order.
struct SymBcoo {
    int nb = 0;         // block rows
    int K = 0;          // upper blocks
    DBuf<int> lptr;     // nb + 1 (build scratch: lower counts)
    DBuf<int2> lref;    // <= K   (build scratch: lower references)
    DBuf<unsigned> lkey, lkey2;  // K (build scratch: column sort keys)
    DBuf<int> lidx;              // K (build scratch: block ids in column order)
    DBuf<char> ltmp;             // radix sort scratch
    DBuf<int> eptr;     // nb + 1
    DBuf<int2> ent;     // lower references + K (<= 2K)
    DBuf<float> valf;   // 12 * K FP32 copy of the BCOO values, padded columns (mixed precision)
    bool valf_ok = false;
    // row-contiguous FP32 copy of both triangles in entry order (mixed
    // precision): entry e of ent as its block, transposed where it is a
    // transposed reference, 10 words (9 column-major floats + its column)
    DBuf<float> valfull;
    bool full_ok = false;
};

// Sort of the static (FEM + mass) triplet segment, built where it started at
// s0 and reused (ids shifted) while its ns row/col indices are unchanged
// (crow/ccol hold the copy checked against on every convert).
struct ConvertPlan {
    bool valid = false;
    long long s0 = 0, ns = 0;
    int nbr = 0, Ks = 0;
    DBuf<int> crow, ccol;       // ns: cached indices
    DBuf<int> start;            // nbr + 1 row starts into sorted
    DBuf<int> sorted;           // ns: triplet ids sorted by (row, col, id)
    DBuf<int> ustart;           // unique-block row starts (nbr + 1)
    DBuf<int> ucol, urun;       // per unique block: column, run start in sorted
};

struct MasState {
    bool ready = false;
    int bank = 16;
    int n = 0;           // FEM vertices covered
    int node_offset = 0; // subtracted from contact-pair vertex ids
    int M = 0;           // slot count (bank * parts)
    int L = 0;           // levels
    int total = 0;       // clusters over all levels (multiple of bank)
    int nbanks = 0;      // total / bank
    std::vector<int> level_size_h;  // 2*(L+1)
    DBuf<int> part_map, real_map;
    DBuf<unsigned> nstart, nnum, nlist;
    DBuf<unsigned> fine_mask, next_mask;
    DBuf<int> prefix0, psum0, next_prefix, next_psum;
    DBuf<int> level_size;  // 2*(L+2)
    DBuf<int> coarse_space;  // L*n
    DBuf<int> going_next;    // cap
    DBuf<unsigned> comp_mask; // cap: bank lanes of each coarse node's component (levels >= 1)
    DBuf<unsigned> proj_key, proj_rec;  // ordered projection: records (entry, block/level), double-buffered
    DBuf<char> proj_sort_tmp;
    DBuf<double> proj_contrib;  // ordered projection: per-record contributions, then piece sums
    size_t going_cap = 0;
    DBuf<int> coarse_table;  // 6*n (itable layout)
    DBuf<int> node_of;       // L*n : global node id of vertex at level l (l=0: slot)
    DBuf<double> bank_upper; // nbanks * npack * 9 (Galerkin, packed upper, col-major blocks)
    DBuf<double> inv64;      // nbanks * D(D+1)/2 symmetric inverse, upper triangle packed by rows (FP64)
    DBuf<float> inv32;       // same, FP32
    bool inv64_ok = true;    // inv64 holds the FP64 inverses (not after an FP32-only build)
    DBuf<unsigned short> inv16;  // half storage: S = D^-1 inv D^-1 as FP16 bits, packed like inv64
    DBuf<float> inv16_d;     // its D: sqrt of the inverse's diagonal, nbanks * 48
    bool inv16_ok = false;   // inv16 built (16-vertex banks, half storage chosen)
    DBuf<int> l1lo, l1hi;    // per level-1 bank: range [lo, hi) of level-0 banks holding its children
    DBuf<int> own_lo, own_hi;  // per level-1 bank: owned vertex range (slot-ordered vertices)
    DBuf<double> proj_scratch; // per-workgroup partials of the top-level bank matrices
    bool slot_ordered = false; // vertex ids increase with their slot
    int nb1 = 0;             // level-1 banks
    DBuf<double> R64, Z64;   // 3*total
    DBuf<float> R32, Z32;
    DBuf<int> sub_index;     // BCOO block indices in the subsystem range
};

struct PcgGraph {
    hipGraphExec_t exec = nullptr;      // a group of two iterations
    hipGraphExec_t exec_big = nullptr;  // a longer group (single-system PCG), same key
    size_t key = 0;
    PcgGraph() = default;
    PcgGraph(const PcgGraph&) = delete;
    PcgGraph& operator=(const PcgGraph&) = delete;
    ~PcgGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (exec_big) (void)hipGraphExecDestroy(exec_big);
    }
};

// Linear BVH over surface primitives (lbvh.hip)
struct LbvhTree {
    int m = 0;
    DBuf<char> leaf, scene, nodes, nbox, sort_tmp;  // boxes / nodes as raw bytes (types private to lbvh.hip)
    DBuf<unsigned> code, idx, flags;
    DBuf<unsigned long long> key;
};
struct Lbvh {
    LbvhTree tri, edge;
    DBuf<int> cnt, off, err;
};

// Batched PCG over scenes concatenated block-row-wise (pcg_batch.hip)
struct BatchWork {
    int nch = 0;                                    // row chunks (never straddling a scene)
    DBuf<int> chunk_row, chunk_scene, scene_chunk;  // nch + 1, nch, S + 1
    DBuf<double> part_pap, part_rr, part_rz;        // one partial per chunk
    DBuf<double> bb, rz, rz0, alpha, beta, tol, segout;  // per scene
    DBuf<int> stop, iters, noconv;                  // per scene
    DBuf<int> ctrl, bad, rstart;                    // C_* words; coupling check counter; row starts
    PcgGraph graph64, graph32, graph_mixed;
    // persistent solver (pcg_persist.hip)
    DBuf<int> order, queue, refine;                 // scenes largest first; scene queue; refinements
    DBuf<int> mlo, mhi, node_scene, bank_scene;     // per scene and MAS level node ranges; scene of each node / bank
    DBuf<double> relres;                            // per-scene true relative residuals
    PcgGraph graph_persist;                         // the fill + persistent launch, captured
};

// Affine bodies coupled to the FEM system (abd.hip)
struct AbdState {
    bool ready = false, inv_ready = false;
    int nb = 0, nav = 0;      // bodies; body vertices (vertex ids 0 .. nav)
    DBuf<int> vstart, body_of, fixed;  // nb + 1 vertex ranges; nav body ids; nb fixed flags
    DBuf<double> xbar, mass, vol;      // 3 nav rest coordinates; 144 nb mass matrices (row-major); nb volumes
    DBuf<double> inv;                  // 144 nb block-Jacobi inverses (row-major)
    DBuf<int> cnt, off, total;         // contact mapping counts / offsets / total
    DBuf<unsigned long long> key;      // vertex-pair keys (double-buffered for the sort)
    DBuf<int> idx;                     // source block ids (double-buffered)
    DBuf<char> sort_tmp;
    DBuf<double> uval;                 // merged vertex-space blocks
    DBuf<int> urow, ucol;
};

// Pinned, device-mapped host words the PCG publishes its progress to, so the
// host follows the solve without copies or stream synchronisation.
struct PinnedPoll {
    int* h = nullptr;  // host view: [0] stop iteration, [1] iteration, [2] publication count; double at [8]
    int* d = nullptr;  // device view of the same memory
    bool failed = false;
    PinnedPoll() = default;
    PinnedPoll(const PinnedPoll&) = delete;
    PinnedPoll& operator=(const PinnedPoll&) = delete;
    ~PinnedPoll() {
        if (h) (void)hipHostFree(h);
    }
};

struct PcgWork {
    DBuf<double> r, z, p, Ap, xw;   // 3N each
    PinnedPoll poll;
    DBuf<double> r64b, d64;         // mixed precision: FP64 residual and correction
    PcgGraph graph64, graph32, graph_mixed;
    DBuf<double> partials;          // per-workgroup partial dot products
    DBuf<double> scal;              // device scalars (see pcg.hip); S_NUM+1 entries used
    DBuf<int> ctrl;                 // done flag, iteration counters, tickets
    DBuf<float> rf, zf, pf, Apf;    // mixed-precision inner vectors
    DBuf<double> Dinv;              // 9 * nb block-Jacobi inverses
};

// Optional HIP-event timing of selected launches (enabled by the caller for
// measurement runs; off by default).
enum ProfSlot { PROF_SPMV = 0, PROF_MAS_SOLVE, PROF_FEM, PROF_CONVERT, PROF_PRECOND_BUILD, PROF_SOLVE, PROF_BARRIER,
                PROF_NSLOTS };
struct Prof {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pend[PROF_NSLOTS];
    double ms[PROF_NSLOTS] = {0};
    long long calls[PROF_NSLOTS] = {0};
};

struct Ctx {
    int device = 0;
    Prof prof;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string last_error;

    // caller-owned triplet store (GIPCTripletMatrix)
    double* t_vals = nullptr;
    int* t_rows = nullptr;
    int* t_cols = nullptr;
    long long t_cap = 0;

    // FEM mesh incidence for the deterministic gradient gather
    int fem_n = 0, fem_T = 0;
    DBuf<int> inc_start;  // n+1
    DBuf<int> inc_list;   // 4T entries: 4*t + corner, ascending t per vertex
    DBuf<double> tet_grad; // 12*T
    int fem_project = 1;   // 1: PSD-projected SNK1 Hessian (Newton); 0: exact (adjoint solves)

    // converted system
    int K = 0;            // unique upper blocks
    int nb = 0;           // block rows of the system
    bool have_bcoo = false;
    DBuf<double> bcoo_val; // 9K scratch before copy-back
    DBuf<int> bcoo_row, bcoo_col;
    DBuf<int> row_count, row_start, uniq_start;
    DBuf<int> dyn_ustart, dyn_ucol, dyn_urun, merge_row;  // contact segment runs, merged rows
    DBuf<int4> merge_desc;                                 // per merged block: contact / static runs
    DBuf<unsigned long long> seg_key, seg_key2;            // (row, col) sort keys of a segment, sorted
    DBuf<int> seg_head, seg_hscan, seg_beg, seg_fin;       // unique-key heads and their scan; row runs
    DBuf<char> seg_tmp;                                    // radix sort / scan scratch
    DBuf<double> energy_part;                              // energy partials (+ total)
    long long static_begin = -1;  // first triplet of the FEM + mass segment (set by FEM assembly)
    ConvertPlan cplan;
    DBuf<int> dev_int;    // small device ints (totals)
    SymBcoo bsr;
    bool have_bsr = false;
    int spmv_group = 8;  // lanes per block row in the PCG SpMV (env SGIPC_SPMV_G: 4, 8, 16)
    int spmv_blocks = 0; // cap on the mixed PCG SpMV's workgroups: 2 per CU (env SGIPC_SPMV_BLOCKS; 0: one per chunk, at most 1024)
    int spmv_full32 = 1; // mixed precision: row-contiguous FP32 copy of both triangles (env SGIPC_SPMV_FULL32=0: upper BCOO)
    int deterministic = 0;  // 1: MAS projection and barrier gradient summed in a fixed order (bitwise repeatable)
    DBuf<double> grad_scr;               // ordered barrier gradient: per-pair rows (12 doubles per pair)
    DBuf<unsigned> grad_key, grad_rec;   // (vertex, record) pairs, double-buffered for the sort
    DBuf<char> grad_sort_tmp;

    // preconditioner
    int precond_kind = 0;
    int bank_inverse = 1;   // MAS bank inverses: BANK_INV_* (default: FP64 MFMA)
    int bank_store = 0;     // storage the fused FP32 bank solves read: BANK_STORE_* (default: FP32)
    int mas_fused = 1;      // top restriction beside the bank solves (16-vertex banks; env SGIPC_MAS_FUSED=0: own launch)
    int sub_offset = 0, sub_nb = 0;
    MasState mas;

    PcgWork pcg;
    AbdState abd;
    BatchWork batch;
    Lbvh lbvh;
    int batch_iters_per_graph = 4;  // batched PCG iterations captured per graph launch
    int batch_persistent = 0;       // batched PCG: 1 = one persistent launch (a workgroup per scene, 16-vertex banks)
    int batch_threads = 1024;       // its workgroup size (512 or 1024)
    int batch_max_wg = 0;           // cap on its workgroups (0: as many as fit the device at once)
    DBuf<double> tune_x;            // autotuner output scratch
    DBuf<int> scan_tmp;
    int sort_impl = 0;              // SORT_OWN, or SORT_ROCPRIM (env SGIPC_SORT=rocprim) as the cross-check
    double inner_tol = 1e-6;        // mixed precision: floor of the inner solves' relative tolerance (env SGIPC_INNER_TOL)
    DBuf<char> sort_tmp;            // radix sort ping-pong buffers and pass state
    DBuf<int> sort_err;             // sticky look-back failure word
    // contact classification workspaces
    DBuf<int> cls_buf, cls_flag, cls_scan, cls_pos;
};

void prof_begin(Ctx& c, int slot);
void prof_end(Ctx& c, int slot);
void prof_flush(Ctx& c);
void prof_release(Ctx& c);
void prof_drop_tail(Ctx& c, int slot, size_t keep);

void exclusive_scan(Ctx& c, const int* in, int* out, int n, int* total_dev, hipStream_t s);
// Stable radix sort of (key, 32-bit value) pairs on the low `bits` bits of
// the keys (radix.hip).  vin == nullptr: the values are id_base + i.  kin
// and kout must differ; values of any 32-bit type move bit for bit.  A
// failed look-back (never expected) sets *err, or the context's sticky
// sort error word when err is null.
enum { SORT_OWN = 0, SORT_ROCPRIM = 1 };
void radix_sort_pairs(Ctx& c, const unsigned long long* kin, unsigned long long* kout, const int* vin, int id_base,
                      int* vout, long long n, int bits, hipStream_t s, int* err = nullptr);
void radix_sort_pairs(Ctx& c, const unsigned* kin, unsigned* kout, const int* vin, int id_base, int* vout,
                      long long n, int bits, hipStream_t s, int* err = nullptr);
int radix_sort_error(Ctx& c);  // the sticky word (synchronises)
// one sort by the library's radix sort or by rocPRIM's (the cross-check); synchronises
void sort_pairs_dispatch(Ctx& c, const void* kin, void* kout, const int* vin, int* vout, long long n, int key_bytes,
                         int bits, int impl);

// component entry points (implemented in the .hip files)
void fem_set_mesh(Ctx& c, int n, int T, const uint32_t* h_tets);
void fem_kinetic_gradient(Ctx& c, int n, const double* x, const double* xt, const double* mass,
                          double* shape);
void fem_assemble(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv,
                  const double* vol, const double* mu, const double* lam, double dt,
                  const int* btype, int index_offset, long long triplet_offset, double* shape);
void fem_mass_triplets(Ctx& c, int n, const double* mass, int index_offset,
                       long long triplet_offset);
void fem_rhs(Ctx& c, int n, const double* fb, const double* shape, const int* btype, double* b);
void fem_param_adjoint(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                       double dt, const int* btype, const double* adj, double* dmu, double* dlam);
int ground_detect(Ctx& c, int m, const unsigned* cand, const double* x, const double* normal, double offset,
                  double dhat, unsigned* out);
void ground_assemble(Ctx& c, int np, const unsigned* gp, const double* x, const double* normal, double offset,
                     double dhat, double kappa, const int* btype, int index_offset, long long triplet_offset,
                     double* fb);
double ground_energy(Ctx& c, int m, const unsigned* ids, const double* x, const double* normal, double offset,
                     double dhat, double kappa);
double ground_step_bound(Ctx& c, int m, const unsigned* ids, const double* x, const double* dir,
                         const double* normal, double offset, double slack);

int bcoo_convert(Ctx& c, long long nt, int nblockrows);
void bsr_build(Ctx& c);
const float* bsr_values_f32(Ctx& c);
const float* bsr_full_f32(Ctx& c);
void spmv64(Ctx& c, double a, const double* x, double b, double* y);

double kinetic_energy(Ctx& c, int n, const double* x, const double* xt, const double* m);
double fem_energy(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                  const double* mu, const double* lam, double dt);
double barrier_energy(Ctx& c, int np, const int* pairs, const double* x, const double* xr, double dhat,
                      double kappa);
// allocation-free SNK1 energy: partials in a static device array (looked up
// once per device by energy_scratch_setup), result to device memory
void energy_scratch_setup(int device);
void fem_energy_noalloc(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                        const double* mu, const double* lam, double dt, double* out);
// energy reductions: workgroups for n items (256 threads), partial buffer,
// fixed-order sum of g partials to the host
int energy_grid(long long n);
// line-search step x = x_temp - alpha dir (free vertices, or all) and the
// largest |dir| over listed vertices (the reference's CFL speed)
void step_forward(Ctx& c, int n, double* x, const double* xt, const double* dir, const int* btype, double alpha,
                  int move_boundary);
double max_speed(Ctx& c, int m, const uint32_t* ids, const double* dir);
double* energy_partials(Ctx& c);
double energy_finish(Ctx& c, double* part, int g);
void mas_set_topology(Ctx& c, int n, int node_offset, const unsigned* nstart, const unsigned* nnum,
                      const unsigned* nlist, int nlist_len, const int* part_map, int M,
                      const int* real_map, int bank);
void mas_build(Ctx& c, int sub_offset, int sub_nb, const int* pairs, int npairs);
// MAS bank inverses: scalar FP64 Gauss-Jordan on the vector ALUs, blocked
// FP64 Gauss-Jordan on MFMA, or blocked FP32 on MFMA + one refinement step
// with the residual accumulated in FP64 (FP32 inverses only)
enum { BANK_INV_VALU64 = 0, BANK_INV_MFMA64 = 1, BANK_INV_MFMA32_REFINED = 2 };
enum { BANK_STORE_F32 = 0, BANK_STORE_F16 = 1 };
// inv16/inv16_d (optional, 16-vertex banks with MFMA64): also write the
// half-storage form of the inverses in the same launch
void bank_inverse_mfma(Ctx& c, int bank, int nbanks, const double* bank_upper, double* inv64, float* inv32,
                       int mode, unsigned short* inv16 = nullptr, float* inv16_d = nullptr);
void mas_apply64(Ctx& c, const double* r, double* z);
void mas_apply32(Ctx& c, const double* r, double* z);
// fused multilevel restriction of r[sub3 .. sub3 + 3n) into R_l (FP64 or FP32
// preconditioner data); bank solves of all levels; both skip when ctrl[0] != 0
// Batched solves: banks whose nodes all belong to one stopped scene are
// skipped (bank_scene: the scene of every bank's nodes, -1 when they belong
// to several or none; sstop: per-scene stop words).  Default: no skipping.
struct SceneSkip {
    const int* bank_scene = nullptr;
    const int* sstop = nullptr;
    __device__ bool stopped(int bank) const {
        if (!bank_scene) return false;
        const int s = bank_scene[bank];
        return s >= 0 && sstop[s] != 0;
    }
};
void mas_restrict(Ctx& c, const double* r, long long sub3, bool fp32, const int* ctrl, bool top = true,
                  SceneSkip skip = SceneSkip{});
void mas_restrict(Ctx& c, const float* r, long long sub3, bool fp32, const int* ctrl, bool top = true,
                  SceneSkip skip = SceneSkip{});
void mas_bank_solve(Ctx& c, bool fp32, const int* ctrl, SceneSkip skip = SceneSkip{});
// PCG iteration with 16-vertex banks: the top-level restriction (levels
// 2 -> L-1) runs in the bank-solve launch, beside the banks of levels 0..2,
// instead of its own launch before it (the x/r update then restricts levels
// 0..2 only: mas_update_restrict(..., top = false))
bool mas_fused(const Ctx& c);
void mas_solve_with_top(Ctx& c, bool fp32, const int* ctrl, SceneSkip skip = SceneSkip{});

// PCG x/r update fused with the restriction of the new residual (levels
// 0..L-1); g1 + gx workgroups, each writing one r.r partial
void mas_update_layout(const Ctx& c, int& g1, int& gx);
void mas_update_restrict(Ctx& c, double* x, double* r, const double* p, const double* Ap, const double* part_pap,
                         int np_pap, const double* part_rz, int np_rz, double* part_rr, int* ctrl, bool fp32,
                         bool top = true);
void mas_update_restrict(Ctx& c, double* x, float* r, const float* p, const float* Ap, const double* part_pap,
                         int np_pap, const double* part_rz, int np_rz, double* part_rr, int* ctrl, bool fp32,
                         bool top = true);
void diag_build(Ctx& c);

void barrier_assemble(Ctx& c, int np, const int* pairs, const int* mat_index, const int counts[3], const double* x,
                      const double* xrest, double dhat, double kappa, const int* btype, int index_offset,
                      long long triplet_offset, double* fb);
int classify_contacts(Ctx& c, int npt, const int* pt, int nee, const int* ee, const double* x, const double* xrest,
                      double dhat, int* pairs, int* mat_index, int capacity, int counts[3]);

struct SolveOpts {
    int precision = 0;  // 0 fp64, 1 mixed
    int stop_rule = 0;  // 0 reference |rz|<=tol rz0, 1 ||r||/||b||<=tol
    double tol = 1e-4;
    double max_iter_ratio = 0.3;
    int max_iter = 0;
    int check_every = 8;
    double inner_tol = 1e-6;  // mixed precision: floor of the inner relative tolerance
};
struct SolveStats {
    int iterations = 0;
    double rel_residual = 0;
    double rz0 = 0;
    int converged = 0;
    int refinements = 0;
};
void pcg_solve(Ctx& c, const double* b, double* x, const SolveOpts& o, SolveStats& st);
void pcg_solve_batched(Ctx& c, int S, const int* h_row_start, const double* b, double* x, const SolveOpts& o,
                       SolveStats* st);
void batch_check_coupling(Ctx& c, const std::vector<int>& row_start);
// the batched solve as one persistent launch: each workgroup takes scenes
// (largest first) and runs a scene's whole solve (validated arguments)
// scene of every MAS node and bank (BatchWork node_scene, bank_scene) and,
// when lo is given, every scene's node range [lo, hi) at every level (S x L)
void batch_scene_maps(Ctx& c, int S, const int* d_row_start, int* lo, int* hi);
void pcg_solve_batched_persistent(Ctx& c, int S, const std::vector<int>& row_start, const double* b, double* x,
                                  const SolveOpts& o, SolveStats* st);
void autotune(Ctx& c, const double* b, int S, const int* h_row_start, int precision, int& group, int& per_graph);
double ccd_step_bound(Ctx& c, int npt, const int* pt, int nee, const int* ee, const double* x, const double* dir,
                      double slack);
// candidate counts in n_pt / n_ee; quadruples written only when they fit the capacities
// (dir non-null: boxes swept from x to x - alpha dir, the CCD candidates)
void broad_phase(Ctx& c, int nsv, const unsigned* sv, int nf, const unsigned* faces, int ne, const unsigned* edges,
                 const double* x, const int* btype, const int* body, double dhat, int* pt_out, int pt_cap, int& n_pt,
                 int* ee_out, int ee_cap, int& n_ee, const double* dir = nullptr, double alpha = 0.0);
void precond_apply(Ctx& c, const double* r, double* z);

void abd_setup(Ctx& c, int nb, int nav, const int* vstart, const double* xbar, const double* mass12,
               const double* vol, const int* fixed);
void abd_positions(Ctx& c, const double* q, double* x);
void abd_assemble(Ctx& c, const double* q, const double* qt, double kappa, double dt, long long toff, double* grad_q);
void abd_contact_gradient(Ctx& c, const double* g, double* grad_q);
long long abd_map_contacts(Ctx& c, long long src, long long m, long long dst);
double abd_energy(Ctx& c, const double* q, const double* qt, double kappa, double dt);
void abd_jacobi_build(Ctx& c);
void abd_precond_apply(Ctx& c, const double* r, double* z);

}  // namespace sgipc
