// Batched PCG + MAS as one persistent launch (replayed from a captured
// graph): every workgroup takes scenes
// from a queue (largest first) and runs a whole scene's solve -- all PCG
// iterations, the MAS apply of every level, and for mixed precision the
// outer FP64 refinement -- with workgroup barriers only.  Scenes share no
// block (checked), and the union MAS is block-diagonal per scene, so no
// workgroup ever waits for another.
//
// Per scene the arithmetic is the reference loop (PCGSolver::pcg,
// S/linear_system/solver/pcg_solver.cu:174-267) as in the launch-based
// batched solver: x0 = 0, stop rule 0 on the pre-update |rz| <= tol rz0 or
// rule 1 on ||r|| <= tol ||b||, iteration budget max_iter, the MAS apply
// z = sum_l R_l^T D_l^-1 R_l r of MASPreconditioner::preconditioning
// (S/MASPreconditioner.cu:2221-2262) restricted to the scene's nodes.  Mixed
// precision: FP32 values, vectors and MAS with FP64 dots; an outer loop
// recomputes the FP64 true residual b - A x and restarts the FP32 solve with
// the tolerance max(0.9 tol ||b|| / ||r||, inner_tol) until
// ||b - A x|| <= tol ||b|| (the single-system mixed solver's rule).
//
// A bank of a coarse level can hold nodes of two scenes; the scene's
// workgroup reads R only of its own nodes (others read as 0: the bank matrix
// couples no two scenes, so its inverse does not either) and writes Z only
// of its own nodes.  All sums have a fixed order (workgroup reduction trees,
// level order), so results repeat bitwise.
#include "spmv_sym.hpp"
#include "pcg_dev.hpp"
#include <algorithm>
#include <numeric>
#include <type_traits>

namespace sgipc {
namespace {

typedef __attribute__((address_space(1))) const void* pg_gptr_t;
typedef __attribute__((address_space(3))) void* pg_lptr_t;

constexpr int kSolveWaves = 8;  // wavefronts that solve banks (one LDS staging area each)

__device__ inline float bcast_lane(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ inline double bcast_lane(double v, int j) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline unsigned pg_ltmask(int lane) { return lane >= 32 ? 0xffffffffu : ((1u << lane) - 1u); }

// component sum of the bank-aligned group in LDS (the restriction's sum):
// the lowest lane of the component sums its lanes in ascending order
template <typename T>
__device__ inline bool comp_sum(unsigned mask, const T* a0, const T* a1, const T* a2, T& s0, T& s1, T& s2) {
    const int t = threadIdx.x, lane = t % 16, base = t - lane;
    if (mask & pg_ltmask(lane)) return false;
    s0 = 0; s1 = 0; s2 = 0;
    for (unsigned m = mask; m; m &= m - 1u) {
        const int l = __ffs(m) - 1;
        s0 += a0[base + l]; s1 += a1[base + l]; s2 += a2[base + l];
    }
    return true;
}

template <typename V, typename MV, typename T>
struct PArgs {
    int S;
    const int* order;        // scenes in processing order (largest first)
    const int* srow;         // S + 1 block-row starts
    const int* mlo;          // S * L: first node of the scene at each MAS level
    const int* mhi;          // S * L: one past its last node
    const int* node_scene;   // scene of every MAS node (-1: empty)
    int* queue;              // next scene to take
    // matrix: entry lists over the upper BCOO; values of the inner solve and FP64 values
    const int* eptr;
    const int2* ent;
    const MV* val;
    const double* val64;
    // MAS over the union (16-vertex banks)
    int L, n;
    const int* part_map;
    const int* going_next;
    const unsigned* fine_mask;
    const unsigned* comp_mask;
    const int* node_of;
    const T* inv;
    T* R;
    T* Z;
    // vectors (union indexing)
    const double* b;
    double* x;
    V* r;
    V* z;
    V* p;
    V* Ap;
    double* r64;  // FP64 residual (mixed)
    double* d;    // FP64 iterate of the inner solve (mixed: the correction)
    // options
    int rule, maxit, mixed;
    double tol, inner_tol;
    // per-scene results
    int* iters;
    int* noconv;
    int* refine;
    double* relres;
};

template <int NTH, int G, typename V, typename MV, typename T>
struct SceneSolver {
    const PArgs<V, MV, T>& a;
    int s, r0, r1;
    double* sh;   // NTH / 64 doubles
    T (*sv)[NTH];
    void* stage;  // kSolveWaves staging areas

    __device__ double block_sum(double v) { return block_sum_all(v, sh); }

    // Ap = A p on the scene's rows, returns p.Ap
    __device__ double spmv_pap() {
        constexpr int RPB = NTH / G;
        const int sub = threadIdx.x % G;
        double acc = 0.0;
        auto pv = [this](size_t i) { return a.p[i]; };
        for (int rb = r0; rb < r1; rb += RPB) {
            const int row = rb + threadIdx.x / G;
            MV s0 = 0, s1 = 0, s2 = 0;
            if (row < r1) sym_row_lane<G, MV>(row, sub, a.eptr, a.ent, a.val, pv, s0, s1, s2);
            group_sum3<G, MV>(s0, s1, s2);
            if (sub < 3 && row < r1) {
                const size_t o = 3 * (size_t)row + sub;
                const V w = (V)(sub == 0 ? s0 : (sub == 1 ? s1 : s2));
                a.Ap[o] = w;
                acc += (double)a.p[o] * (double)w;
            }
        }
        __threadfence_block();
        return block_sum(acc);
    }

    // r64 = b - A x in FP64 on the scene's rows, returns its squared norm
    __device__ double true_residual() {
        constexpr int RPB = NTH / G;
        const int sub = threadIdx.x % G;
        double acc = 0.0;
        auto xv = [this](size_t i) { return a.x[i]; };
        for (int rb = r0; rb < r1; rb += RPB) {
            const int row = rb + threadIdx.x / G;
            double s0 = 0, s1 = 0, s2 = 0;
            if (row < r1) sym_row_lane<G, double>(row, sub, a.eptr, a.ent, a.val64, xv, s0, s1, s2);
            group_sum3<G, double>(s0, s1, s2);
            if (sub < 3 && row < r1) {
                const size_t o = 3 * (size_t)row + sub;
                const double rv = a.b[o] - (sub == 0 ? s0 : (sub == 1 ? s1 : s2));
                a.r64[o] = rv;
                acc += rv * rv;
            }
        }
        __threadfence_block();
        return block_sum(acc);
    }

    // one bank of 48 rows by one wavefront, R of the scene's own nodes only
    __device__ void bank_solve(int bank, void* area) {
        constexpr int D = 48, NP = D * (D + 1) / 2;
        using V4 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
        constexpr int NV = NP * (int)sizeof(T) / 16;
        constexpr int PER = (NV + 63) / 64;
        V4* As = static_cast<V4*>(area);
        const int i = threadIdx.x & 63;
        const int node = bank * 16 + min(i, D - 1) / 3;
        const bool own = i < D && a.node_scene[node] == s;
        const T rv = own ? a.R[(size_t)bank * D + i] : (T)0;
        const V4* src = reinterpret_cast<const V4*>(a.inv + (size_t)bank * NP);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int q0 = k * 64;
            __builtin_amdgcn_global_load_lds((pg_gptr_t)(src + min(q0 + i, NV - 1)), (pg_lptr_t)(As + q0), 16, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        const T* A = reinterpret_cast<const T*>(As);
        T acc = 0;
        if (i < D) {
            int idx = i;
#pragma unroll 8
            for (int j = 0; j < D; ++j) {
                acc += A[idx] * bcast_lane(rv, j);
                idx += (j < i) ? (D - 1 - j) : 1;
            }
        }
        if (own) a.Z[(size_t)bank * D + i] = acc;
        __builtin_amdgcn_wave_barrier();
    }

    // z = M^-1 r on the scene (r holds the residual), returns r.z
    __device__ double mas_apply() {
        const int t = threadIdx.x, L = a.L;
        const int* lo = a.mlo + (size_t)s * L;
        const int* hi = a.mhi + (size_t)s * L;
        // level 0 -> 1: the scene's level-0 banks (slots); R_0 and R_1
        {
            const int c_beg = lo[0] / 16 * 16, c_end = (hi[0] + 15) / 16 * 16;
            for (int c0 = c_beg; c0 < c_end; c0 += NTH) {
                const int slot = c0 + t;
                const int v0i = slot < c_end ? a.part_map[slot] : -1;
                const int v = (v0i >= r0 && v0i < r1) ? v0i : -1;  // the scene's vertices only
                T v0 = 0, v1 = 0, v2 = 0;
                if (v >= 0) {
                    const V* rv = a.r + 3 * (size_t)v;
                    v0 = (T)rv[0]; v1 = (T)rv[1]; v2 = (T)rv[2];
                    T* Rs = a.R + 3 * (size_t)slot;
                    Rs[0] = v0; Rs[1] = v1; Rs[2] = v2;
                }
                sv[0][t] = v0; sv[1][t] = v1; sv[2][t] = v2;
                __syncthreads();
                T s0, s1, s2;
                if (v >= 0 && L > 1 && comp_sum<T>(a.fine_mask[v], sv[0], sv[1], sv[2], s0, s1, s2)) {
                    T* Rq = a.R + 3 * (size_t)a.going_next[v];
                    Rq[0] = s0; Rq[1] = s1; Rq[2] = s2;
                }
                __syncthreads();
            }
        }
        __threadfence_block();
        __syncthreads();
        // level l -> l + 1 on the scene's nodes of level l
        for (int l = 1; l + 1 < L; ++l) {
            const int nlo = lo[l], nhi = hi[l];
            const int c_beg = nlo / 16 * 16, c_end = (nhi + 15) / 16 * 16;
            for (int c0 = c_beg; c0 < c_end; c0 += NTH) {
                const int g = c0 + t;
                const bool own = g >= nlo && g < nhi && a.node_scene[g] == s;
                T v0 = 0, v1 = 0, v2 = 0;
                if (own) { v0 = a.R[3 * (size_t)g]; v1 = a.R[3 * (size_t)g + 1]; v2 = a.R[3 * (size_t)g + 2]; }
                sv[0][t] = v0; sv[1][t] = v1; sv[2][t] = v2;
                __syncthreads();
                T s0, s1, s2;
                if (own && comp_sum<T>(a.comp_mask[g], sv[0], sv[1], sv[2], s0, s1, s2)) {
                    T* Rq = a.R + 3 * (size_t)a.going_next[g];
                    Rq[0] = s0; Rq[1] = s1; Rq[2] = s2;
                }
                __syncthreads();
            }
            __threadfence_block();
            __syncthreads();
        }
        // bank solves of every level (the banks holding the scene's nodes)
        const int wv = t >> 6;
        if (wv < kSolveWaves) {
            using V4 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
            constexpr int PER = ((48 * 49 / 2) * (int)sizeof(T) / 16 + 63) / 64;
            void* area = static_cast<V4*>(stage) + (size_t)wv * PER * 64;
            int k = wv;
            for (int l = 0; l < L; ++l) {
                const int b0 = lo[l] / 16, b1 = (hi[l] + 15) / 16;  // banks [b0, b1)
                const int nbl = hi[l] > lo[l] ? b1 - b0 : 0;
                for (; k < nbl; k += kSolveWaves) bank_solve(b0 + k, area);
                k -= nbl;
            }
        }
        __threadfence_block();
        __syncthreads();
        // collect: z = sum_l Z_l[node_l(v)] in level order, and r.z
        double acc = 0.0;
        for (int row = r0 + t; row < r1; row += NTH) {
            T s0 = 0, s1 = 0, s2 = 0;
            for (int l = 0; l < L; ++l) {
                const size_t g = 3 * (size_t)a.node_of[(size_t)l * a.n + row];
                s0 += a.Z[g]; s1 += a.Z[g + 1]; s2 += a.Z[g + 2];
            }
            const size_t o = 3 * (size_t)row;
            const V z0 = (V)s0, z1 = (V)s1, z2 = (V)s2;
            a.z[o] = z0; a.z[o + 1] = z1; a.z[o + 2] = z2;
            acc += (double)a.r[o] * (double)z0 + (double)a.r[o + 1] * (double)z1 + (double)a.r[o + 2] * (double)z2;
        }
        __threadfence_block();
        return block_sum(acc);
    }

    // PCG on the scene with r holding the right-hand side and the iterate in
    // xi (zeroed here); returns the iteration count (the reference's), sets
    // noconv when the budget stopped it
    __device__ int pcg(double* xi, double tol, int rule, int budget, bool& noconv) {
        const int t = threadIdx.x;
        const long long i0 = 3LL * r0, i1 = 3LL * r1;
        double bbp = 0.0;
        for (long long i = i0 + t; i < i1; i += NTH) {
            xi[i] = 0.0;
            bbp += (double)a.r[i] * (double)a.r[i];
        }
        const double bb = block_sum(bbp);
        double rz = mas_apply();
        const double rz0 = rz;
        noconv = false;
        if (rz == 0.0 || bb == 0.0) return 0;  // zero right-hand side: x = 0 is exact
        if (budget <= 1) { noconv = true; return budget; }
        for (long long i = i0 + t; i < i1; i += NTH) a.p[i] = a.z[i];
        __threadfence_block();
        __syncthreads();
        for (int k = 1;; ++k) {
            const double pap = spmv_pap();
            const double al = rz / pap;
            double rrp = 0.0;
            for (long long i = i0 + t; i < i1; i += NTH) {
                xi[i] += al * (double)a.p[i];
                const V ri = (V)((double)a.r[i] - al * (double)a.Ap[i]);
                a.r[i] = ri;
                rrp += (double)ri * (double)ri;
            }
            __threadfence_block();
            const double rr = block_sum(rrp);
            const double rzn = mas_apply();
            const bool stop = rule == 0 ? fabs(rz) <= tol * rz0 : sqrt(rr) <= tol * sqrt(bb);
            if (stop) return k;
            if (k + 1 >= budget) { noconv = true; return budget; }
            const double be = rzn / rz;
            rz = rzn;
            for (long long i = i0 + t; i < i1; i += NTH) a.p[i] = (V)((double)a.z[i] + be * (double)a.p[i]);
            __threadfence_block();
            __syncthreads();
        }
    }

    __device__ void solve() {
        const int t = threadIdx.x;
        const long long i0 = 3LL * r0, i1 = 3LL * r1;
        int total = 0, refinements = 0;
        bool conv = false, nc = false;
        double rel = 0.0;
        if (!a.mixed) {
            for (long long i = i0 + t; i < i1; i += NTH) a.r[i] = (V)a.b[i];
            __threadfence_block();
            __syncthreads();
            total = pcg(a.x, a.tol, a.rule, a.maxit, nc);
            conv = !nc;
            const double rr = true_residual();
            double bp = 0.0;
            for (long long i = i0 + t; i < i1; i += NTH) bp += a.b[i] * a.b[i];
            const double bb = block_sum(bp);
            rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
        } else {
            double bp = 0.0;
            for (long long i = i0 + t; i < i1; i += NTH) {
                a.x[i] = 0.0;
                a.r64[i] = a.b[i];
                bp += a.b[i] * a.b[i];
            }
            __threadfence_block();
            const double bb = block_sum(bp);
            const double bn = sqrt(bb);
            double rr = bb;
            for (int outer = 0; outer < 50; ++outer) {
                const double rn = sqrt(rr);
                if (rn <= a.tol * bn || bn == 0.0) break;
                if (total >= a.maxit) { nc = true; break; }
                const double eta = fmax(0.9 * a.tol * bn / rn, a.inner_tol);
                for (long long i = i0 + t; i < i1; i += NTH) a.r[i] = (V)a.r64[i];
                __threadfence_block();
                __syncthreads();
                bool inc = false;
                total += pcg(a.d, eta, 1, a.maxit - total, inc);
                refinements = outer + 1;
                for (long long i = i0 + t; i < i1; i += NTH) a.x[i] += a.d[i];
                __threadfence_block();
                __syncthreads();
                rr = true_residual();
            }
            conv = !nc && (sqrt(rr) <= a.tol * bn || bn == 0.0);
            rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
        }
        if (t == 0) {
            a.iters[s] = total;
            a.noconv[s] = conv ? 0 : 1;
            a.refine[s] = refinements;
            a.relres[s] = rel;
        }
    }
};

template <int NTH, int G, typename V, typename MV, typename T>
__global__ void __launch_bounds__(NTH) k_pcg_persist(PArgs<V, MV, T> a) {
    using V4 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    constexpr int PER = ((48 * 49 / 2) * (int)sizeof(T) / 16 + 63) / 64;
    __shared__ double sh[NTH / 64];
    __shared__ T sv[3][NTH];
    __shared__ V4 stage[kSolveWaves * PER * 64];
    __shared__ int take;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) take = atomicAdd(a.queue, 1);
        __syncthreads();
        const int q = take;
        if (q >= a.S) return;
        const int s = a.order[q];
        SceneSolver<NTH, G, V, MV, T> so{a, s, a.srow[s], a.srow[s + 1], sh, sv, stage};
        so.solve();
    }
}

// node range [lo, hi) of every scene at every MAS level, and the scene of
// every node (a node's vertices are connected, so they share one scene)
__global__ void __launch_bounds__(256) k_scene_levels(int n, int L, int S, const int* __restrict__ srow,
                                                       const int* __restrict__ node_of, int* __restrict__ lo,
                                                       int* __restrict__ hi, int* __restrict__ node_scene) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = v < n;  // no early exit: every lane takes part in the wavefront reductions
    int a = -1;
    if (valid) {
        int a1 = 0, b = S;  // largest s with srow[s] <= v (srow[S] = n > v)
        while (b - a1 > 1) {
            const int m = (a1 + b) >> 1;
            if (srow[m] <= v) a1 = m; else b = m;
        }
        a = a1;
    }
    // a wavefront's vertices are consecutive, so they nearly always share one
    // scene: then lane 0 (the lowest vertex) updates the range with the
    // wavefront's min / max, else every lane updates it itself
    const int a0 = __shfl(a, 0, 64);
    const bool uniform = __all(!valid || a == a0);
    for (int l = 0; l < L; ++l) {
        const int g = valid ? node_of[(size_t)l * n + v] : 0;
        if (valid) node_scene[g] = a;
        if (!lo) continue;
        if (uniform) {
            int gmin = valid ? g : 0x7fffffff, gmax = valid ? g : -1;
            for (int off = 32; off > 0; off >>= 1) {
                gmin = min(gmin, __shfl_xor(gmin, off, 64));
                gmax = max(gmax, __shfl_xor(gmax, off, 64));
            }
            if ((threadIdx.x & 63) == 0 && a0 >= 0) {
                atomicMin(lo + (size_t)a0 * L + l, gmin);
                atomicMax(hi + (size_t)a0 * L + l, gmax + 1);
            }
        } else if (valid) {
            atomicMin(lo + (size_t)a * L + l, g);
            atomicMax(hi + (size_t)a * L + l, g + 1);
        }
    }
}

// scene of every bank's nodes: the one scene they all belong to, or -1
// (nodes of several scenes, or none)
__global__ void k_bank_scene(int nbanks, int B, const int* __restrict__ node_scene, int* __restrict__ bank_scene) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbanks) return;
    int sc = -2;
    for (int i = 0; i < B; ++i) {
        const int q = node_scene[(size_t)b * B + i];
        if (q < 0) continue;
        sc = (sc == -2 || sc == q) ? q : -1;
    }
    bank_scene[b] = sc == -2 ? -1 : sc;
}

template <int NTH, int G, typename V, typename MV, typename T>
int occupancy_of() {
    int n = 0;
    SGIPC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pcg_persist<NTH, G, V, MV, T>, NTH, 0));
    return n;
}

template <int NTH, int G, typename V, typename MV, typename T>
int persist_grid(Ctx& c, int S) {
    int ncu = 0;
    SGIPC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device));
    const int per_cu = std::max(1, occupancy_of<NTH, G, V, MV, T>());
    int grid = std::max(1, std::min(S, per_cu * std::max(ncu, 1)));
    if (c.batch_max_wg > 0) grid = std::min(grid, c.batch_max_wg);
    return grid;
}

template <typename V, typename MV, typename T>
size_t persist_key(const PArgs<V, MV, T>& a, int G, int NTH, int grid, hipStream_t s) {
    size_t h = 1469598103934665603ull;
    auto mix = [&](size_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
    for (const void* q : {(const void*)a.order, (const void*)a.srow, (const void*)a.mlo, (const void*)a.mhi,
                          (const void*)a.node_scene, (const void*)a.queue, (const void*)a.eptr, (const void*)a.ent,
                          (const void*)a.val, (const void*)a.val64, (const void*)a.part_map,
                          (const void*)a.going_next, (const void*)a.fine_mask, (const void*)a.comp_mask,
                          (const void*)a.node_of, (const void*)a.inv, (const void*)a.R, (const void*)a.Z,
                          (const void*)a.b, (const void*)a.x, (const void*)a.r, (const void*)a.z, (const void*)a.p,
                          (const void*)a.Ap, (const void*)a.r64, (const void*)a.d, (const void*)a.iters,
                          (const void*)a.noconv, (const void*)a.refine, (const void*)a.relres, (const void*)s})
        mix((size_t)q);
    for (int v : {a.S, a.L, a.n, a.rule, a.maxit, a.mixed, G, NTH, grid, (int)sizeof(V), (int)sizeof(MV),
                  (int)sizeof(T)})
        mix((size_t)(unsigned)v);
    mix(std::hash<double>{}(a.tol));
    mix(std::hash<double>{}(a.inner_tol));
    return h;
}

// The queue reset (a fill kernel, no memset) and the persistent launch,
// captured once into a graph per argument set and replayed.
template <typename V, typename MV, typename T>
void run_persist(Ctx& c, const PArgs<V, MV, T>& a, int S) {
    const int G = c.spmv_group, NTH = c.batch_threads;
    hipStream_t s = c.stream;
    int grid = 0;
    auto pick = [&](auto nth, auto body) {
        constexpr int N = decltype(nth)::value;
        if (G == 4) body(std::integral_constant<int, N>{}, std::integral_constant<int, 4>{});
        else if (G == 16) body(std::integral_constant<int, N>{}, std::integral_constant<int, 16>{});
        else body(std::integral_constant<int, N>{}, std::integral_constant<int, 8>{});
    };
    auto each = [&](auto body) {
        if (NTH == 512) pick(std::integral_constant<int, 512>{}, body);
        else pick(std::integral_constant<int, 1024>{}, body);
    };
    each([&](auto nth, auto g) { grid = persist_grid<decltype(nth)::value, decltype(g)::value, V, MV, T>(c, S); });
    const size_t key = persist_key(a, G, NTH, grid, s);
    PcgGraph& gc = c.batch.graph_persist;
    if (!gc.exec || gc.key != key) {
        if (gc.exec) { (void)hipGraphExecDestroy(gc.exec); gc.exec = nullptr; }
        hipGraph_t gr = nullptr;
        SGIPC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            fill<int>(a.queue, 0, 1, s);
            each([&](auto nth, auto g) {
                hipLaunchKernelGGL((k_pcg_persist<decltype(nth)::value, decltype(g)::value, V, MV, T>), dim3(grid),
                                   dim3(decltype(nth)::value), 0, s, a);
            });
            SGIPC_LAUNCH_CHECK();
        } catch (...) {
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(s, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            throw;
        }
        SGIPC_HIP(hipStreamEndCapture(s, &gr));
        hipError_t e = hipGraphInstantiate(&gc.exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        SGIPC_HIP(e);
        gc.key = key;
    }
    SGIPC_HIP(hipGraphLaunch(gc.exec, s));
}

}  // namespace

void batch_scene_maps(Ctx& c, int S, const int* d_rs, int* lo, int* hi) {
    MasState& m = c.mas;
    BatchWork& w = c.batch;
    hipStream_t s = c.stream;
    const int L = m.L;
    int* nsc = w.node_scene.reserve((size_t)m.total + 1);
    int* bsc = w.bank_scene.reserve((size_t)m.nbanks + 1);
    if (lo) {
        fill<int>(lo, 0x7fffffff, (long long)S * L, s);
        fill<int>(hi, 0, (long long)S * L, s);
    }
    fill<int>(nsc, -1, (long long)m.total, s);
    if (m.n > 0) {
        hipLaunchKernelGGL(k_scene_levels, dim3(div_up(m.n, 256)), dim3(256), 0, s, m.n, L, S, d_rs,
                           (const int*)m.node_of.p, lo, hi, nsc);
        SGIPC_LAUNCH_CHECK();
    }
    if (m.nbanks > 0) {
        hipLaunchKernelGGL(k_bank_scene, dim3(div_up(m.nbanks, 256)), dim3(256), 0, s, m.nbanks, m.bank,
                           (const int*)nsc, bsc);
        SGIPC_LAUNCH_CHECK();
    }
}

void pcg_solve_batched_persistent(Ctx& c, int S, const std::vector<int>& rs, const double* b, double* x,
                                  const SolveOpts& o, SolveStats* st) {
    MasState& m = c.mas;
    if (m.bank != 16) throw state_error("solve_batched: the persistent solver needs 16-vertex banks");
    if (c.batch_threads != 512 && c.batch_threads != 1024)
        throw arg_error("solve_batched: persistent workgroup size must be 512 or 1024");
    hipStream_t s = c.stream;
    BatchWork& w = c.batch;
    const long long N = 3LL * c.nb;
    const int L = m.L;
    // scene descriptors: row starts, order (largest first), node ranges per level
    int* d_rs = w.rstart.reserve((size_t)S + 1);
    SGIPC_HIP(hipMemcpyAsync(d_rs, rs.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice, s));
    std::vector<int> ord(S);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int i, int j) { return rs[i + 1] - rs[i] > rs[j + 1] - rs[j]; });
    int* d_ord = w.order.reserve((size_t)S);
    SGIPC_HIP(hipMemcpyAsync(d_ord, ord.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
    int* lo = w.mlo.reserve((size_t)S * L + 1);
    int* hi = w.mhi.reserve((size_t)S * L + 1);
    batch_scene_maps(c, S, d_rs, lo, hi);
    const int* nsc = w.node_scene.p;
    for (DBuf<int>* q : {&w.iters, &w.noconv, &w.refine, &w.queue}) q->reserve((size_t)S + 1);
    w.relres.reserve((size_t)S + 1);
    PcgWork& pw = c.pcg;
    const int max_iter = o.max_iter > 0 ? o.max_iter : (int)(o.max_iter_ratio * (double)N);
    auto common = [&](auto& a) {
        a.S = S; a.order = d_ord; a.srow = d_rs; a.mlo = lo; a.mhi = hi; a.node_scene = nsc; a.queue = w.queue.p;
        a.eptr = c.bsr.eptr.p; a.ent = c.bsr.ent.p; a.val64 = c.t_vals;
        a.L = L; a.n = m.n; a.part_map = m.part_map.p; a.going_next = m.going_next.p; a.fine_mask = m.fine_mask.p;
        a.comp_mask = m.comp_mask.p; a.node_of = m.node_of.p;
        a.b = b; a.x = x; a.r64 = pw.r64b.reserve(N + 1); a.d = pw.d64.reserve(N + 1);
        a.maxit = max_iter; a.tol = o.tol; a.inner_tol = o.inner_tol;
        a.iters = w.iters.p; a.noconv = w.noconv.p; a.refine = w.refine.p; a.relres = w.relres.p;
    };
    if (o.precision == 0) {
        auto run64 = [&](auto tag) {
            using T = decltype(tag);
            PArgs<double, double, T> a{};
            common(a);
            a.val = c.t_vals;
            a.inv = sizeof(T) == 4 ? (const T*)(const void*)m.inv32.p : (const T*)(const void*)m.inv64.p;
            a.R = sizeof(T) == 4 ? (T*)(void*)m.R32.p : (T*)(void*)m.R64.p;
            a.Z = sizeof(T) == 4 ? (T*)(void*)m.Z32.p : (T*)(void*)m.Z64.p;
            a.r = pw.r.reserve(N + 1); a.z = pw.z.reserve(N + 1); a.p = pw.p.reserve(N + 1);
            a.Ap = pw.Ap.reserve(N + 1);
            a.rule = o.stop_rule; a.mixed = 0;
            run_persist(c, a, S);
        };
        if (c.precond_kind == 3) run64(float{});
        else {
            if (!m.inv64_ok) throw state_error("solve_batched: only FP32 inverses were built");
            run64(double{});
        }
    } else {
        if (!m.inv32.p) throw state_error("solve_batched: mixed precision needs the FP32 MAS inverses");
        PArgs<float, float, float> a{};
        common(a);
        a.val = bsr_values_f32(c);
        a.inv = m.inv32.p; a.R = m.R32.p; a.Z = m.Z32.p;
        a.r = pw.rf.reserve(N + 1); a.z = pw.zf.reserve(N + 1); a.p = pw.pf.reserve(N + 1);
        a.Ap = pw.Apf.reserve(N + 1);
        a.rule = 1; a.mixed = 1;
        run_persist(c, a, S);
    }
    std::vector<int> hit(S), hnc(S), hrf(S);
    std::vector<double> hrel(S);
    SGIPC_HIP(hipMemcpyAsync(hit.data(), w.iters.p, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(hnc.data(), w.noconv.p, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(hrf.data(), w.refine.p, sizeof(int) * S, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(hrel.data(), w.relres.p, sizeof(double) * S, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < S; ++i) {
        st[i] = SolveStats{};
        st[i].iterations = hit[i];
        st[i].converged = hnc[i] ? 0 : 1;
        st[i].refinements = hrf[i];
        st[i].rel_residual = hrel[i];
    }
}

}  // namespace sgipc
