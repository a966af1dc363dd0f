// Batched PCG: many independent scenes concatenated into one converted
// system (scene s owns block rows [row_start[s], row_start[s+1]); no block
// couples two scenes), each scene running its own preconditioned conjugate
// gradients -- its own rz, alpha, beta and stop decision -- in the same
// kernel launches.  The MAS preconditioner is built once over the union: its
// banks aggregate connected nodes only, so it is block-diagonal per scene
// and each scene sees exactly its own preconditioner.
//
// Layout: block rows are cut into chunks that never straddle a scene (1024/G
// rows for the SpMV with G lanes per row); one workgroup per chunk writes one
// partial per dot product, and a scalar kernel (one wave per scene) sums
// each scene's partials in chunk order (deterministic, no FP atomics) into
// per-scene alpha / beta / stop words.  A stopped scene's rows are left untouched by
// every later kernel.  The iteration sequence
//   SpMV (Ap, p.Ap) -> alpha -> x, r update (r.r) -> MAS restrict + bank solve
//   -> collect (z, r.z) -> beta + stop tests -> p update
// is captured once into a HIP graph (several iterations per graph) and
// replayed until every scene has stopped.
//
// Per scene the arithmetic is the reference loop (PCGSolver::pcg,
// S/linear_system/solver/pcg_solver.cu:174-267): x0 = 0, stop rule 0 on the
// pre-update |rz| <= tol rz0, or rule 1 on ||r|| <= tol ||b||, iteration
// count as the reference's.  The reference has no batched solver.
#include "spmv_sym.hpp"
#include "pcg_dev.hpp"
#include <cmath>

namespace sgipc {
namespace {

constexpr int kBT = 256;       // threads of the row-wise batched kernels
constexpr int kScalT = 1024;   // the scalar kernel: one wave per scene, 16 scenes per workgroup

enum BPhase { B_INIT_BB = 0, B_INIT_RZ, B_ALPHA, B_BETA, B_SEGSUM };

__device__ inline double wave_sum_partials(const double* __restrict__ part, int lo, int hi) {
    const int lane = threadIdx.x & 63;
    double v = 0.0;
    for (int i = lo + lane; i < hi; i += 64) v += part[i];
    return wave_sum(v);
}

// Ap = A p on the chunk's rows, partial p.Ap per chunk
template <int G, typename MV, typename V>
__global__ void __launch_bounds__(1024) k_bspmv(const int* __restrict__ chunk_row, const int* __restrict__ chunk_scene,
                                                const int* __restrict__ eptr, const int2* __restrict__ ent,
                                                const MV* __restrict__ val, const V* __restrict__ p,
                                                V* __restrict__ Ap, const int* __restrict__ sstop,
                                                double* __restrict__ part, const int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[1024 / 64];
    const int ch = blockIdx.x;
    if (sstop[chunk_scene[ch]]) {
        if (threadIdx.x == 0) part[ch] = 0.0;
        return;
    }
    const int r0 = chunk_row[ch], r1 = chunk_row[ch + 1];
    const int row = r0 + threadIdx.x / G, sub = threadIdx.x % G;
    auto pv = [p](size_t i) { return p[i]; };
    MV s0 = 0, s1 = 0, s2 = 0;
    if (row < r1) sym_row_lane<G, MV>(row, sub, eptr, ent, val, pv, s0, s1, s2);
    group_sum3<G, MV>(s0, s1, s2);
    double contrib = 0.0;
    if (sub < 3 && row < r1) {
        const size_t o = 3 * (size_t)row + sub;
        const V a = (V)(sub == 0 ? s0 : (sub == 1 ? s1 : s2));
        Ap[o] = a;
        contrib = (double)p[o] * (double)a;
    }
    write_partial(contrib, part, sh);
}

// The same over the row-contiguous FP32 copy of both triangles (mixed
// solves): the G lanes of a row stream consecutive entries; same entry
// order and products as k_bspmv, so the same sums.
template <int G>
__global__ void __launch_bounds__(1024) k_bspmv_full(const int* __restrict__ chunk_row,
                                                     const int* __restrict__ chunk_scene, const int* __restrict__ eptr,
                                                     const float* __restrict__ vf,
                                                     const float* __restrict__ p, float* __restrict__ Ap,
                                                     const int* __restrict__ sstop, double* __restrict__ part,
                                                     const int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[1024 / 64];
    const int ch = blockIdx.x;
    if (sstop[chunk_scene[ch]]) {
        if (threadIdx.x == 0) part[ch] = 0.0;
        return;
    }
    const int r0 = chunk_row[ch], r1 = chunk_row[ch + 1];
    const int row = r0 + threadIdx.x / G, sub = threadIdx.x % G;
    float s0 = 0, s1 = 0, s2 = 0;
    if (row < r1) {
        const int end = eptr[row + 1];
        int e = eptr[row] + sub;
        for (; e + G < end; e += 2 * G) {  // two entries in flight
            const FullEntry fa = load_full_entry(vf, e), fb = load_full_entry(vf, e + G);
            const f3_a4 xa = load_z3(p, fa.col), xb = load_z3(p, fb.col);
            full_madd(fa, xa, s0, s1, s2);
            full_madd(fb, xb, s0, s1, s2);
        }
        if (e < end) {
            const FullEntry f = load_full_entry(vf, e);
            full_madd(f, load_z3(p, f.col), s0, s1, s2);
        }
    }
    group_sum3<G, float>(s0, s1, s2);
    double contrib = 0.0;
    if (sub < 3 && row < r1) {
        const size_t o = 3 * (size_t)row + sub;
        const float a = sub == 0 ? s0 : (sub == 1 ? s1 : s2);
        Ap[o] = a;
        contrib = (double)p[o] * (double)a;
    }
    write_partial(contrib, part, sh);
}

// x += alpha_s p ; r -= alpha_s Ap ; partial r.r per chunk
template <typename V>
__global__ void __launch_bounds__(kBT) k_bupdate(const int* __restrict__ chunk_row, const int* __restrict__ chunk_scene,
                                                 double* __restrict__ x, V* __restrict__ r, const V* __restrict__ p,
                                                 const V* __restrict__ Ap, const double* __restrict__ alpha,
                                                 const int* __restrict__ sstop, double* __restrict__ part,
                                                 const int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[kBT / 64];
    const int ch = blockIdx.x, s = chunk_scene[ch];
    if (sstop[s]) {
        if (threadIdx.x == 0) part[ch] = 0.0;
        return;
    }
    const double al = alpha[s];
    const long long i1 = 3LL * chunk_row[ch + 1];
    double v = 0.0;
    for (long long i = 3LL * chunk_row[ch] + threadIdx.x; i < i1; i += kBT) {
        x[i] += al * (double)p[i];
        V ri = (V)((double)r[i] - al * (double)Ap[i]);
        r[i] = ri;
        v += (double)ri * (double)ri;
    }
    write_partial(v, part, sh);
}

// z = sum_l Z_l[node_l(v)] (MAS over all rows), partial r.z per chunk
template <typename V, typename T>
__global__ void __launch_bounds__(kBT) k_bcollect(const int* __restrict__ chunk_row, const int* __restrict__ chunk_scene,
                                                  const V* __restrict__ r, V* __restrict__ z, int n, int L,
                                                  const int* __restrict__ node_of, const T* __restrict__ Z,
                                                  const int* __restrict__ sstop, double* __restrict__ part,
                                                  const int* __restrict__ ctrl) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[kBT / 64];
    const int ch = blockIdx.x;
    if (sstop[chunk_scene[ch]]) {
        if (threadIdx.x == 0) part[ch] = 0.0;
        return;
    }
    double v = 0.0;
    for (int row = chunk_row[ch] + threadIdx.x; row < chunk_row[ch + 1]; row += kBT) {
        T s0 = 0, s1 = 0, s2 = 0;
        for (int l = 0; l < L; ++l) {
            size_t g = 3 * (size_t)node_of[(size_t)l * n + row];
            s0 += Z[g]; s1 += Z[g + 1]; s2 += Z[g + 2];
        }
        const size_t o = 3 * (size_t)row;
        const V z0 = (V)s0, z1 = (V)s1, z2 = (V)s2;
        z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
        v += (double)r[o] * (double)z0 + (double)r[o + 1] * (double)z1 + (double)r[o + 2] * (double)z2;
    }
    write_partial(v, part, sh);
}

// p = z + beta_s p (beta_s = 0 on the first iteration)
template <typename V>
__global__ void __launch_bounds__(kBT) k_bpupdate(const int* __restrict__ chunk_row, const int* __restrict__ chunk_scene,
                                                  const V* __restrict__ z, V* __restrict__ p,
                                                  const double* __restrict__ beta, const int* __restrict__ sstop,
                                                  const int* __restrict__ ctrl, int first) {
    if (!first && ctrl[C_STOP]) return;
    const int ch = blockIdx.x, s = chunk_scene[ch];
    if (sstop[s]) return;
    const double be = first ? 0.0 : beta[s];
    const long long i1 = 3LL * chunk_row[ch + 1];
    for (long long i = 3LL * chunk_row[ch] + threadIdx.x; i < i1; i += kBT)
        p[i] = first ? z[i] : (V)((double)z[i] + be * (double)p[i]);
}

// partials of a.b per chunk (all scenes)
template <typename VA, typename VB>
__global__ void __launch_bounds__(kBT) k_bdot(const int* __restrict__ chunk_row, const VA* __restrict__ a,
                                              const VB* __restrict__ b, double* __restrict__ part) {
    __shared__ double sh[kBT / 64];
    const int ch = blockIdx.x;
    const long long i1 = 3LL * chunk_row[ch + 1];
    double v = 0.0;
    for (long long i = 3LL * chunk_row[ch] + threadIdx.x; i < i1; i += kBT) v += (double)a[i] * (double)b[i];
    write_partial(v, part, sh);
}

struct BScal {
    double *bb, *rz, *rz0, *alpha, *beta, *tol, *out;
    int *stop, *iters, *noconv;
};

// extra control words of the batched solver (after the shared C_* words)
constexpr int kCActive = C_NUM;      // active scenes counted by the workgroups of a scalar kernel
constexpr int kCTicket = C_NUM + 1;  // workgroups of that kernel done

// Per-scene scalars: one wave per scene, 16 scenes per workgroup.  The
// phases that decide the stop (INIT_RZ, BETA) count active scenes per
// workgroup; the last workgroup to finish sets the stop word and advances
// the iteration counter (every workgroup has read both before it finishes).
__global__ void __launch_bounds__(kScalT) k_bscalars(int phase, int S, const int* __restrict__ scene_chunk,
                                                     const double* __restrict__ pa, const double* __restrict__ pb,
                                                     BScal q, int rule, int* __restrict__ ctrl) {
    if ((phase == B_ALPHA || phase == B_BETA) && ctrl[C_STOP]) return;
    __shared__ int act[kScalT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = ctrl[C_ITER];
    int active = 0;
    for (int s = blockIdx.x * (kScalT / 64) + w; s < S && s < (blockIdx.x + 1) * (kScalT / 64); s += kScalT / 64) {
        const int lo = scene_chunk[s], hi = scene_chunk[s + 1];
        if (phase == B_SEGSUM) {
            double v = wave_sum_partials(pa, lo, hi);
            if (lane == 0) q.out[s] = v;
            continue;
        }
        if (phase == B_INIT_BB) {
            double v = wave_sum_partials(pa, lo, hi);
            if (lane == 0) q.bb[s] = v;
            continue;
        }
        if (phase == B_INIT_RZ) {
            double v = wave_sum_partials(pa, lo, hi);
            if (lane == 0) {
                q.rz[s] = v;
                q.rz0[s] = v;
                q.iters[s] = 0;
                q.noconv[s] = 0;
                if (v == 0.0 || q.bb[s] == 0.0) q.stop[s] = 1;  // zero right-hand side: x = 0 is exact
                if (!q.stop[s] && ctrl[C_MAXIT] <= 1) {        // no iteration fits the budget
                    q.stop[s] = 1;
                    q.iters[s] = ctrl[C_MAXIT];
                    q.noconv[s] = 1;
                }
                if (!q.stop[s]) active += 1;
            }
            continue;
        }
        if (q.stop[s]) continue;
        if (phase == B_ALPHA) {
            double pap = wave_sum_partials(pa, lo, hi);
            if (lane == 0) q.alpha[s] = q.rz[s] / pap;
            continue;
        }
        // B_BETA: pa = r.r partials (updated r), pb = r.z partials (new z)
        double rr = wave_sum_partials(pa, lo, hi);
        double rzn = wave_sum_partials(pb, lo, hi);
        if (lane == 0) {
            bool stop = rule == 0 ? fabs(q.rz[s]) <= q.tol[s] * q.rz0[s] : sqrt(rr) <= q.tol[s] * sqrt(q.bb[s]);
            if (stop) {
                q.stop[s] = 1;
                q.iters[s] = k;
            } else if (k + 1 >= ctrl[C_MAXIT]) {
                q.stop[s] = 1;
                q.iters[s] = ctrl[C_MAXIT];
                q.noconv[s] = 1;
            } else {
                q.beta[s] = rzn / q.rz[s];
                q.rz[s] = rzn;
                active += 1;
            }
        }
    }
    if (phase == B_SEGSUM || phase == B_INIT_BB || phase == B_ALPHA) return;
    active = (int)wave_sum((double)active);
    if (lane == 0) act[w] = active;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < kScalT / 64; ++i) tot += act[i];
        if (tot) atomicAdd(ctrl + kCActive, tot);
        __threadfence();
        if (atomicAdd(ctrl + kCTicket, 1) == (int)gridDim.x - 1) {  // last workgroup
            __threadfence();
            const int all = atomicAdd(ctrl + kCActive, 0);
            ctrl[kCActive] = 0;
            ctrl[kCTicket] = 0;
            if (phase == B_INIT_RZ) {
                ctrl[C_ITER] = 1;
                ctrl[C_STOP] = all == 0 ? 1 : 0;
                ctrl[C_NOCONV] = 0;
            } else {
                ctrl[C_ITER] = k + 1;
                if (all == 0) ctrl[C_STOP] = 1;
            }
        }
    }
}

__global__ void k_bctrl_init(int* ctrl, int maxit) {
    ctrl[C_STOP] = 0;
    ctrl[C_ITER] = 1;
    ctrl[C_MAXIT] = maxit;
    ctrl[C_NOCONV] = 0;
    ctrl[kCActive] = 0;
    ctrl[kCTicket] = 0;
}

template <typename TD, typename TS>
__global__ void k_bcast(long long N, TD* __restrict__ dst, const TS* __restrict__ src) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        dst[i] = (TD)src[i];
}

__global__ void k_baxpy(long long N, double* __restrict__ y, const double* __restrict__ x) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x)
        y[i] += x[i];
}

// counts blocks whose row and column lie in different scenes
__global__ void k_coupling(int K, const int* __restrict__ rows, const int* __restrict__ cols,
                           const int* __restrict__ rs, int S, int* __restrict__ bad) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    auto scene_of = [&](int v) {
        int lo = 0, hi = S;  // largest s with rs[s] <= v
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (rs[mid] <= v) lo = mid; else hi = mid;
        }
        return lo;
    };
    if (scene_of(rows[k]) != scene_of(cols[k])) atomicAdd(bad, 1);
}

int flat_grid(long long N) {
    long long g = div_up(N, 256);
    return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

template <typename V, typename MV, typename T>
struct BEngine {
    Ctx& c;
    BatchWork& w;
    long long N;
    int S, nch, G;
    V *r, *z, *p, *Ap;
    double* x;
    const MV* val;
    bool fp32mas;
    int rule;
    const float* vfull = nullptr;  // mixed: row-contiguous FP32 copy of both triangles (else the upper BCOO)

    BScal scal() {
        return BScal{w.bb.p, w.rz.p, w.rz0.p, w.alpha.p, w.beta.p, w.tol.p, w.segout.p, w.stop.p, w.iters.p,
                     w.noconv.p};
    }
    T* Rm() { return fp32mas ? (T*)(void*)c.mas.R32.p : (T*)(void*)c.mas.R64.p; }
    T* Zm() { return fp32mas ? (T*)(void*)c.mas.Z32.p : (T*)(void*)c.mas.Z64.p; }

    void scalars(int phase, const double* pa, const double* pb) {
        hipLaunchKernelGGL(k_bscalars, dim3(div_up(S, kScalT / 64)), dim3(kScalT), 0, c.stream, phase, S,
                           (const int*)w.scene_chunk.p, pa, pb, scal(), rule, w.ctrl.p);
        SGIPC_LAUNCH_CHECK();
    }

    void precond() {
        // z = M^-1 r over the union, then r.z partials per chunk
        const SceneSkip skip{w.bank_scene.p, w.stop.p};
        const bool fused = mas_fused(c);  // top-level restriction inside the bank-solve launch
        mas_restrict(c, r, 0, fp32mas, w.ctrl.p, !fused, skip);
        if (fused) mas_solve_with_top(c, fp32mas, w.ctrl.p, skip);
        else mas_bank_solve(c, fp32mas, w.ctrl.p, skip);
        hipLaunchKernelGGL((k_bcollect<V, T>), dim3(nch), dim3(kBT), 0, c.stream, (const int*)w.chunk_row.p,
                           (const int*)w.chunk_scene.p, (const V*)r, z, c.mas.n, c.mas.L, (const int*)c.mas.node_of.p,
                           (const T*)Zm(), (const int*)w.stop.p, w.part_rz.p, (const int*)w.ctrl.p);
        SGIPC_LAUNCH_CHECK();
    }

    void iteration() {
        hipStream_t s = c.stream;
        const SymBcoo& B = c.bsr;
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(nch), dim3(1024), 0, s, (const int*)w.chunk_row.p,
                               (const int*)w.chunk_scene.p, (const int*)B.eptr.p, (const int2*)B.ent.p, val,
                               (const V*)p, Ap, (const int*)w.stop.p, w.part_pap.p, (const int*)w.ctrl.p);
        };
        if constexpr (sizeof(MV) == 4 && sizeof(V) == 4) {
            if (vfull) {
                auto launchf = [&](auto kern) {
                    hipLaunchKernelGGL(kern, dim3(nch), dim3(1024), 0, s, (const int*)w.chunk_row.p,
                                       (const int*)w.chunk_scene.p, (const int*)B.eptr.p, vfull,
                                       (const float*)p, (float*)Ap, (const int*)w.stop.p, w.part_pap.p,
                                       (const int*)w.ctrl.p);
                };
                if (G == 4) launchf(k_bspmv_full<4>);
                else if (G == 16) launchf(k_bspmv_full<16>);
                else launchf(k_bspmv_full<8>);
            }
        }
        if (!vfull) {
            if (G == 4) launch(k_bspmv<4, MV, V>);
            else if (G == 16) launch(k_bspmv<16, MV, V>);
            else launch(k_bspmv<8, MV, V>);
        }
        SGIPC_LAUNCH_CHECK();
        scalars(B_ALPHA, w.part_pap.p, nullptr);
        hipLaunchKernelGGL((k_bupdate<V>), dim3(nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p,
                           (const int*)w.chunk_scene.p, x, r, (const V*)p, (const V*)Ap, (const double*)w.alpha.p,
                           (const int*)w.stop.p, w.part_rr.p, (const int*)w.ctrl.p);
        SGIPC_LAUNCH_CHECK();
        precond();
        scalars(B_BETA, w.part_rr.p, w.part_rz.p);
        hipLaunchKernelGGL((k_bpupdate<V>), dim3(nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p,
                           (const int*)w.chunk_scene.p, (const V*)z, p, (const double*)w.beta.p,
                           (const int*)w.stop.p, (const int*)w.ctrl.p, 0);
        SGIPC_LAUNCH_CHECK();
    }

    // r = rhs (already in r), x = 0; per-scene tolerances and initial stop
    // flags are in w.tol / w.stop.  Computes bb, z, rz, rz0, p.
    void init(int maxit) {
        hipStream_t s = c.stream;
        hipLaunchKernelGGL(k_bctrl_init, dim3(1), dim3(1), 0, s, w.ctrl.p, maxit);
        SGIPC_LAUNCH_CHECK();
        fill<double>(x, 0.0, N, s);
        hipLaunchKernelGGL((k_bdot<V, V>), dim3(nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p, (const V*)r,
                           (const V*)r, w.part_rr.p);
        SGIPC_LAUNCH_CHECK();
        scalars(B_INIT_BB, w.part_rr.p, nullptr);
        precond();
        scalars(B_INIT_RZ, w.part_rz.p, nullptr);
        hipLaunchKernelGGL((k_bpupdate<V>), dim3(nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p,
                           (const int*)w.chunk_scene.p, (const V*)z, p, (const double*)w.beta.p,
                           (const int*)w.stop.p, (const int*)w.ctrl.p, 1);
        SGIPC_LAUNCH_CHECK();
    }

    size_t key(int per_graph) const {
        size_t h = 1469598103934665603ull;
        auto mix = [&](size_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
        const SymBcoo& B = c.bsr;
        mix((size_t)N); mix((size_t)S); mix((size_t)nch); mix((size_t)G); mix((size_t)per_graph);
        mix((size_t)rule); mix((size_t)r); mix((size_t)z); mix((size_t)p); mix((size_t)Ap); mix((size_t)x);
        mix((size_t)val); mix((size_t)fp32mas); mix((size_t)B.eptr.p); mix((size_t)B.ent.p); mix((size_t)B.nb);
        mix((size_t)vfull); mix((size_t)mas_fused(c));
        for (const void* q : {(const void*)w.chunk_row.p, (const void*)w.chunk_scene.p, (const void*)w.scene_chunk.p,
                              (const void*)w.part_pap.p, (const void*)w.part_rr.p, (const void*)w.part_rz.p,
                              (const void*)w.bb.p, (const void*)w.rz.p, (const void*)w.rz0.p, (const void*)w.alpha.p,
                              (const void*)w.beta.p, (const void*)w.tol.p, (const void*)w.segout.p,
                              (const void*)w.stop.p, (const void*)w.iters.p, (const void*)w.noconv.p,
                              (const void*)w.ctrl.p, (const void*)w.bank_scene.p})
            mix((size_t)q);
        mix((size_t)c.stream);
        const MasState& m = c.mas;
        mix((size_t)m.nbanks); mix((size_t)m.L); mix((size_t)m.M); mix((size_t)m.n); mix((size_t)m.bank);
        mix((size_t)m.nb1); mix((size_t)m.node_of.p); mix((size_t)m.inv64.p); mix((size_t)m.inv32.p);
        mix((size_t)m.R64.p); mix((size_t)m.R32.p); mix((size_t)m.Z64.p); mix((size_t)m.Z32.p);
        mix((size_t)m.going_next.p); mix((size_t)m.level_size.p); mix((size_t)m.l1lo.p);
        mix((size_t)m.inv16_ok); mix((size_t)m.inv16.p); mix((size_t)m.inv16_d.p);
        return h;
    }

    void run(PcgGraph& gc, int per_graph) {
        hipStream_t s = c.stream;
        int hc[C_NUM] = {0};
        SGIPC_HIP(hipMemcpyAsync(hc, w.ctrl.p, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        if (hc[C_STOP]) return;
        size_t k = key(per_graph);
        if (!gc.exec || gc.key != k) {
            if (gc.exec) { (void)hipGraphExecDestroy(gc.exec); gc.exec = nullptr; }
            hipGraph_t g = nullptr;
            SGIPC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                for (int i = 0; i < per_graph; ++i) iteration();
            } catch (...) {
                hipGraph_t dead = nullptr;
                (void)hipStreamEndCapture(s, &dead);
                if (dead) (void)hipGraphDestroy(dead);
                throw;
            }
            SGIPC_HIP(hipStreamEndCapture(s, &g));
            hipError_t e = hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            SGIPC_HIP(e);
            gc.key = k;
        }
        for (;;) {
            SGIPC_HIP(hipGraphLaunch(gc.exec, s));
            SGIPC_HIP(hipMemcpyAsync(hc, w.ctrl.p, sizeof(int) * C_NUM, hipMemcpyDeviceToHost, s));
            SGIPC_HIP(hipStreamSynchronize(s));
            if (hc[C_STOP]) break;
        }
    }
};

// per-scene sums of a chunk-partial array, to the host
void seg_to_host(Ctx& c, BatchWork& w, int S, const double* part, std::vector<double>& out) {
    hipLaunchKernelGGL(k_bscalars, dim3(div_up(S, kScalT / 64)), dim3(kScalT), 0, c.stream, (int)B_SEGSUM, S,
                       (const int*)w.scene_chunk.p,
                       part, (const double*)nullptr,
                       BScal{w.bb.p, w.rz.p, w.rz0.p, w.alpha.p, w.beta.p, w.tol.p, w.segout.p, w.stop.p, w.iters.p,
                             w.noconv.p},
                       0, w.ctrl.p);
    SGIPC_LAUNCH_CHECK();
    out.resize(S);
    SGIPC_HIP(hipMemcpyAsync(out.data(), w.segout.p, sizeof(double) * S, hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
}

// chunk table: scenes cut into chunks of at most `rows` block rows
void build_chunks(Ctx& c, BatchWork& w, const std::vector<int>& rs, int rows) {
    const int S = (int)rs.size() - 1;
    std::vector<int> crow, cscene, schunk(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        schunk[s] = (int)cscene.size();
        for (int r0 = rs[s]; r0 < rs[s + 1]; r0 += rows) {
            crow.push_back(r0);
            cscene.push_back(s);
        }
    }
    schunk[S] = (int)cscene.size();
    crow.push_back(rs[S]);
    const int nch = (int)cscene.size();
    w.nch = nch;
    w.chunk_row.reserve(nch + 1);
    w.chunk_scene.reserve(nch + 1);
    w.scene_chunk.reserve(S + 1);
    SGIPC_HIP(hipMemcpy(w.chunk_row.p, crow.data(), sizeof(int) * (nch + 1), hipMemcpyHostToDevice));
    if (nch) SGIPC_HIP(hipMemcpy(w.chunk_scene.p, cscene.data(), sizeof(int) * nch, hipMemcpyHostToDevice));
    SGIPC_HIP(hipMemcpy(w.scene_chunk.p, schunk.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice));
    for (DBuf<double>* b : {&w.part_pap, &w.part_rr, &w.part_rz}) b->reserve(nch + 1);
}

}  // namespace

void batch_check_coupling(Ctx& c, const std::vector<int>& rs) {
    BatchWork& w = c.batch;
    const int S = (int)rs.size() - 1;
    int* d_rs = w.rstart.reserve((size_t)S + 1);
    int* bad = w.bad.reserve(1);
    SGIPC_HIP(hipMemcpyAsync(d_rs, rs.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice, c.stream));
    fill<int>(bad, 0, 1, c.stream);
    if (c.K > 0) {
        hipLaunchKernelGGL(k_coupling, dim3(div_up(c.K, 256)), dim3(256), 0, c.stream, c.K, (const int*)c.t_rows,
                           (const int*)c.t_cols, (const int*)d_rs, S, bad);
        SGIPC_LAUNCH_CHECK();
    }
    int h = 0;
    SGIPC_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    if (h) throw arg_error("solve_batched: " + std::to_string(h) + " matrix blocks couple two scenes");
}

void pcg_solve_batched(Ctx& c, int S, const int* h_row_start, const double* b, double* x, const SolveOpts& o,
                       SolveStats* st) {
    if (!c.have_bcoo) throw state_error("solve_batched: convert first");
    const bool mas = (c.precond_kind == 2 || c.precond_kind == 3);
    if (!mas || !c.mas.ready) throw state_error("solve_batched: needs a built MAS preconditioner");
    if (c.sub_offset != 0 || c.mas.n != c.nb)
        throw state_error("solve_batched: the MAS preconditioner must cover every block row");
    if (S <= 0) throw arg_error("solve_batched: no scenes");
    if (o.stop_rule != 0 && o.stop_rule != 1) throw arg_error("solve_batched: stop rule must be 0 or 1");
    std::vector<int> rs(h_row_start, h_row_start + S + 1);
    if (rs[0] != 0 || rs[S] != c.nb) throw arg_error("solve_batched: row_start must run from 0 to the block rows");
    for (int s = 0; s < S; ++s)
        if (rs[s + 1] < rs[s]) throw arg_error("solve_batched: row_start must be non-decreasing");
    if (!c.have_bsr) bsr_build(c);
    if (c.batch_persistent && c.mas.bank == 16) {
        batch_check_coupling(c, rs);
        pcg_solve_batched_persistent(c, S, rs, b, x, o, st);
        return;
    }
    hipStream_t s = c.stream;
    BatchWork& w = c.batch;
    const int G = c.spmv_group;
    build_chunks(c, w, rs, 1024 / G);
    batch_check_coupling(c, rs);
    batch_scene_maps(c, S, w.rstart.p, nullptr, nullptr);  // banks of stopped scenes are skipped
    for (DBuf<double>* q : {&w.bb, &w.rz, &w.rz0, &w.alpha, &w.beta, &w.tol, &w.segout}) q->reserve(S + 1);
    for (DBuf<int>* q : {&w.stop, &w.iters, &w.noconv}) q->reserve(S + 1);
    w.ctrl.reserve(C_NUM + 8);
    const long long N = 3LL * c.nb;
    const int max_iter = o.max_iter > 0 ? o.max_iter : (int)(o.max_iter_ratio * (double)N);
    const int per_graph = std::max(1, c.batch_iters_per_graph);
    std::vector<int> hit(S), hnc(S), hstop(S, 0);
    std::vector<double> htol(S, o.tol);
    PcgWork& pw = c.pcg;
    for (int i = 0; i < S; ++i) st[i] = SolveStats{};
    if (o.precision == 0) {
        auto run64 = [&](auto tag) {
            using T = decltype(tag);
            BEngine<double, double, T> E{c, w, N, S, w.nch, G, pw.r.reserve(N + 1), pw.z.reserve(N + 1),
                                         pw.p.reserve(N + 1), pw.Ap.reserve(N + 1), x, (const double*)c.t_vals,
                                         sizeof(T) == 4, o.stop_rule};
            hipLaunchKernelGGL((k_bcast<double, double>), dim3(flat_grid(N)), dim3(256), 0, s, N, E.r, b);
            SGIPC_LAUNCH_CHECK();
            SGIPC_HIP(hipMemcpyAsync(w.tol.p, htol.data(), sizeof(double) * S, hipMemcpyHostToDevice, s));
            SGIPC_HIP(hipMemcpyAsync(w.stop.p, hstop.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
            E.init(max_iter);
            E.run(sizeof(T) == 4 ? w.graph32 : w.graph64, per_graph);
        };
        if (c.precond_kind == 3) run64(float{});
        else run64(double{});
        SGIPC_HIP(hipMemcpyAsync(hit.data(), w.iters.p, sizeof(int) * S, hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipMemcpyAsync(hnc.data(), w.noconv.p, sizeof(int) * S, hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < S; ++i) {
            st[i].iterations = hit[i];
            st[i].converged = hnc[i] ? 0 : 1;
        }
    } else {
        // FP32 inner solves (FP32 values, FP32 MAS, FP64 dots) with FP64
        // outer refinement on each scene's true residual
        if (c.precond_kind != 3 && c.precond_kind != 2) throw arg_error("solve_batched: unsupported preconditioner");
        if (!c.mas.inv32.p) throw state_error("solve_batched: mixed precision needs the FP32 MAS inverses");
        float* rf = pw.rf.reserve(N + 1);
        double* r64 = pw.r64b.reserve(N + 1);
        double* d64 = pw.d64.reserve(N + 1);
        // the SpMV reads the row-contiguous FP32 copy of both triangles, or
        // (SGIPC_SPMV_FULL32=0) the FP32 copy of the upper BCOO; only one is built
        const float* vfull = c.spmv_full32 ? bsr_full_f32(c) : nullptr;
        BEngine<float, float, float> E{c, w, N, S, w.nch, G, rf, pw.zf.reserve(N + 1), pw.pf.reserve(N + 1),
                                      pw.Apf.reserve(N + 1), d64, vfull ? nullptr : bsr_values_f32(c), true, 1};
        E.vfull = vfull;
        fill<double>(x, 0.0, N, s);
        hipLaunchKernelGGL((k_bcast<double, double>), dim3(flat_grid(N)), dim3(256), 0, s, N, r64, b);
        SGIPC_LAUNCH_CHECK();
        std::vector<double> bb, rr;
        hipLaunchKernelGGL((k_bdot<double, double>), dim3(w.nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p, b, b,
                           w.part_rr.p);
        SGIPC_LAUNCH_CHECK();
        seg_to_host(c, w, S, w.part_rr.p, bb);
        rr = bb;
        std::vector<int> total(S, 0), done(S, 0), nc(S, 0);
        for (int outer = 0; outer < 50; ++outer) {
            int left = 0;
            for (int i = 0; i < S; ++i) {
                const double bn = std::sqrt(bb[i]), rn = std::sqrt(rr[i]);
                if (!done[i] && (rn <= o.tol * bn || bn == 0.0)) done[i] = 1;
                if (!done[i] && total[i] >= max_iter) { done[i] = 1; nc[i] = 1; }
                hstop[i] = done[i];
                htol[i] = done[i] ? 0.0 : std::max(0.9 * o.tol * bn / rn, o.inner_tol);
                left += done[i] ? 0 : 1;
            }
            if (!left) break;
            int budget = 0;
            for (int i = 0; i < S; ++i)
                if (!done[i]) budget = std::max(budget, max_iter - total[i]);
            hipLaunchKernelGGL((k_bcast<float, double>), dim3(flat_grid(N)), dim3(256), 0, s, N, rf,
                               (const double*)r64);
            SGIPC_LAUNCH_CHECK();
            SGIPC_HIP(hipMemcpyAsync(w.tol.p, htol.data(), sizeof(double) * S, hipMemcpyHostToDevice, s));
            SGIPC_HIP(hipMemcpyAsync(w.stop.p, hstop.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
            E.init(budget);
            E.run(w.graph_mixed, per_graph);
            SGIPC_HIP(hipMemcpyAsync(hit.data(), w.iters.p, sizeof(int) * S, hipMemcpyDeviceToHost, s));
            SGIPC_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < S; ++i)
                if (!done[i]) {
                    total[i] += hit[i];
                    st[i].refinements = outer + 1;
                }
            hipLaunchKernelGGL(k_baxpy, dim3(flat_grid(N)), dim3(256), 0, s, N, x, (const double*)d64);
            SGIPC_LAUNCH_CHECK();
            hipLaunchKernelGGL((k_bcast<double, double>), dim3(flat_grid(N)), dim3(256), 0, s, N, r64, b);
            SGIPC_LAUNCH_CHECK();
            spmv64(c, -1.0, x, 1.0, r64);
            hipLaunchKernelGGL((k_bdot<double, double>), dim3(w.nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p,
                               (const double*)r64, (const double*)r64, w.part_rr.p);
            SGIPC_LAUNCH_CHECK();
            seg_to_host(c, w, S, w.part_rr.p, rr);
        }
        for (int i = 0; i < S; ++i) {
            const double bn = std::sqrt(bb[i]);
            st[i].iterations = total[i];
            st[i].converged = (!nc[i] && (std::sqrt(rr[i]) <= o.tol * bn || bn == 0.0)) ? 1 : 0;
        }
    }
    // per-scene FP64 true residuals
    double* tmp = pw.xw.reserve(N + 1);
    hipLaunchKernelGGL((k_bcast<double, double>), dim3(flat_grid(N)), dim3(256), 0, s, N, tmp, b);
    SGIPC_LAUNCH_CHECK();
    spmv64(c, -1.0, x, 1.0, tmp);
    std::vector<double> rr, bb;
    hipLaunchKernelGGL((k_bdot<double, double>), dim3(w.nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p,
                       (const double*)tmp, (const double*)tmp, w.part_rr.p);
    SGIPC_LAUNCH_CHECK();
    seg_to_host(c, w, S, w.part_rr.p, rr);
    hipLaunchKernelGGL((k_bdot<double, double>), dim3(w.nch), dim3(kBT), 0, s, (const int*)w.chunk_row.p, b, b,
                       w.part_rr.p);
    SGIPC_LAUNCH_CHECK();
    seg_to_host(c, w, S, w.part_rr.p, bb);
    for (int i = 0; i < S; ++i) st[i].rel_residual = bb[i] > 0.0 ? std::sqrt(rr[i] / bb[i]) : std::sqrt(rr[i]);
}

}  // namespace sgipc

namespace sgipc {

// Schedule autotuner: times a fixed number of PCG iterations (unreachable
// tolerance) for each candidate schedule on the current system and keeps the
// fastest.  Candidates: lanes per block row in the SpMV (4, 8, 16) and, for
// the batched solver, iterations captured per graph launch (2, 4, 8) of the
// launch sequence, and (16-vertex banks) the persistent form with 512 or
// 1024 threads per workgroup.  For the batched solver the fixed-iteration
// trial only shortlists: the fastest launch sequence and the fastest
// persistent form of each workgroup size are then timed on a whole solve to
// a 1e-6 true relative residual, where scenes stop at their own iteration
// counts (which can rank the persistent forms the other way round), and the
// fastest of those is kept.  Returns the chosen (group, iterations per
// graph); the chosen batched form stays in the context.
void autotune(Ctx& c, const double* b, int S, const int* h_row_start, int precision, int& group, int& per_graph) {
    if (!c.have_bcoo) throw state_error("autotune: convert first");
    const long long N = 3LL * c.nb;
    double* x = c.tune_x.reserve(N + 1);
    SolveOpts o;
    o.precision = precision;
    o.stop_rule = 1;
    o.tol = 0.0;            // never met: every candidate runs the same iteration count
    o.max_iter = 48;
    o.inner_tol = 0.0;
    o.check_every = 48;
    struct Cand {
        int group, per_graph, persistent, threads;
    };
    std::vector<Cand> cands;
    for (int g : {4, 8, 16}) {
        if (S > 0)
            for (int pg : {2, 4, 8}) cands.push_back({g, pg, 0, c.batch_threads});
        else
            cands.push_back({g, c.batch_iters_per_graph, 0, c.batch_threads});
    }
    if (S > 0 && c.mas.ready && c.mas.bank == 16)
        for (int g : {4, 8, 16})
            for (int th : {512, 1024}) cands.push_back({g, c.batch_iters_per_graph, 1, th});
    const Cand save{c.spmv_group, c.batch_iters_per_graph, c.batch_persistent, c.batch_threads};
    auto apply = [&](const Cand& k) {
        c.spmv_group = k.group;
        c.batch_iters_per_graph = k.per_graph;
        c.batch_persistent = k.persistent;
        c.batch_threads = k.threads;
    };
    double best = 1e300;
    Cand pick = save;
    int best_pg = save.per_graph;  // iterations per graph of the fastest launch-sequence candidate
    double best_seq = 1e300;
    // shortlist for the whole-solve round: fastest sequence, fastest
    // persistent form with 512 and with 1024 threads
    Cand short_c[3] = {save, save, save};
    double short_ms[3] = {1e300, 1e300, 1e300};
    std::vector<SolveStats> st(S > 0 ? S : 1);
    hipEvent_t e0, e1;
    SGIPC_HIP(hipEventCreate(&e0));
    SGIPC_HIP(hipEventCreate(&e1));
    auto time_solve = [&](const SolveOpts& so) {
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // first run builds the graphs
            SGIPC_HIP(hipEventRecord(e0, c.stream));
            if (S > 0) pcg_solve_batched(c, S, h_row_start, b, x, so, st.data());
            else pcg_solve(c, b, x, so, st[0]);
            SGIPC_HIP(hipEventRecord(e1, c.stream));
            SGIPC_HIP(hipEventSynchronize(e1));
            SGIPC_HIP(hipEventElapsedTime(&ms, e0, e1));
        }
        return (double)ms;
    };
    try {
        for (const Cand& k : cands) {
            apply(k);
            const double ms = time_solve(o);
            if (ms < best) { best = ms; pick = k; }
            if (!k.persistent && ms < best_seq) { best_seq = ms; best_pg = k.per_graph; }
            const int slot = !k.persistent ? 0 : (k.threads == 512 ? 1 : 2);
            if (ms < short_ms[slot]) { short_ms[slot] = ms; short_c[slot] = k; }
        }
        if (S > 0 && short_ms[1] < 1e300) {
            SolveOpts full;
            full.precision = precision;
            full.stop_rule = 1;
            full.tol = 1e-6;
            full.inner_tol = c.inner_tol;
            double best_full = 1e300;
            for (int i = 0; i < 3; ++i) {
                if (short_ms[i] >= 1e300) continue;
                apply(short_c[i]);
                const double ms = time_solve(full);
                if (ms < best_full) { best_full = ms; pick = short_c[i]; }
            }
        }
    } catch (...) {
        apply(save);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (pick.persistent) pick.per_graph = best_pg;
    apply(pick);
    group = pick.group;
    per_graph = pick.per_graph;
}

}  // namespace sgipc
