// Multilevel additive Schwarz (MAS) preconditioner: aggregation tables,
// Galerkin bank matrices, Gauss-Jordan bank inverses and the apply
//   z = sum_l R_l^T D_l^{-1} R_l r.
//
// Reference: S/MASPreconditioner.cu -- ReorderRealtime :1772-1813 and its
// kernels (:63-627), _buildCollisionConnection_new :1132-1540,
// PrepareHessian_bcoo :1816-2062, __inverse6_P96x96 :630-726,
// preconditioning :2221-2262 (__buildMultiLevelR_optimized_new :729-847,
// _schwarzLocalXSym6 :957-1027, __collectFinalZ_new :850-879),
// computeNumLevels :1724-1743; bank map S/gl_main.cu:1066-1093.
//
// Table semantics are identical to the reference (integer, bit-exact); the
// reference's warp-synchronous 16-lane-bank tricks become explicit LDS stages
// with block barriers (a bank is a quarter of a wave64).  The restriction is
// a deterministic per-bank component sum (no FP atomics); the apply is a
// bank-per-workgroup dense GEMV on the symmetric inverse.
#include "context.hpp"
#include <rocprim/device/device_radix_sort.hpp>
#include <utility>
#include "pcg_dev.hpp"
#include <algorithm>

namespace sgipc {
namespace {

constexpr int kTB = 256;  // threads per block for bank-parallel kernels

__device__ inline unsigned ltmask(int lane) { return lane >= 32 ? 0xffffffffu : ((1u << lane) - 1u); }
__device__ inline int lowbit(unsigned m) { return __ffs(m) - 1; }

// closure of lane's mask over the original masks of its bank (BFS on bits)
__device__ inline unsigned closure(unsigned own, int lane, const unsigned* bank_masks) {
    unsigned conn = own, visited = 1u << lane;
    for (int it = 0; it < 32; ++it) {
        unsigned todo = visited ^ conn;
        if (!todo) break;
        int nv = lowbit(todo);
        visited |= 1u << nv;
        conn |= bank_masks[nv];
    }
    return conn;
}

// ---- level 0 ---------------------------------------------------------------
constexpr int kNbrBatch = 8;  // neighbour-list gathers in flight per thread

__global__ void k_l0_masks(int M, int B, const int* __restrict__ part_map, const int* __restrict__ real_map,
                           const unsigned* __restrict__ nstart, const unsigned* __restrict__ nnum,
                           const unsigned* __restrict__ nlist, unsigned* __restrict__ fine_mask) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    int v = part_map[s];
    if (v < 0) return;
    int bank = s / B, lane = s % B;
    unsigned msk = 1u << lane, st = nstart[v], nn = nnum[v];
    // neighbours eight at a time: their ids, then their slots, are loaded
    // before any is used, so the gathers overlap (the OR is order-free)
    for (unsigned i0 = 0; i0 < nn; i0 += kNbrBatch) {
        int ids[kNbrBatch], su[kNbrBatch];
#pragma unroll
        for (int k = 0; k < kNbrBatch; ++k) ids[k] = (i0 + k < nn) ? (int)nlist[st + i0 + k] : -1;
#pragma unroll
        for (int k = 0; k < kNbrBatch; ++k) su[k] = ids[k] >= 0 ? real_map[ids[k]] : -1;
#pragma unroll
        for (int k = 0; k < kNbrBatch; ++k)
            if (ids[k] >= 0 && su[k] / B == bank) msk |= 1u << (su[k] % B);
    }
    fine_mask[v] = msk;
}

// contact-pair connectivity (decode of the int4 sign encoding)
__device__ inline int decode_pair(int4 q, int node_offset, int* v) {
    int x = q.x, y = q.y, z = q.z, w = q.w, cnt;
    if (x >= 0) {
        if (w < 0) w = -w - 1;
        v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4;
    } else {
        x = -x - 1;
        if (z < 0) {
            if (y < 0) { y = -y - 1; z = -z - 1; w = -w - 1; v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4; }
            else { v[0] = x; v[1] = y; v[2] = -1; v[3] = -1; cnt = 2; }
        } else if (w < 0) {
            if (y < 0) { y = -y - 1; w = -w - 1; v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4; }
            else { v[0] = x; v[1] = y; v[2] = z; v[3] = -1; cnt = 3; }
        } else { v[0] = x; v[1] = y; v[2] = z; v[3] = w; cnt = 4; }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (i < cnt) ? v[i] - node_offset : -1;
    return cnt;
}

// level < 0: ids are slots and dst is indexed by vertex; else ids are the
// level-local coarse ids from coarse_space row (level-1), dst indexed by id.
__global__ void k_collision(int np, const int4* __restrict__ pairs, int node_offset, int n, int B, int level,
                            const int* __restrict__ real_map, const int* __restrict__ coarse_space,
                            unsigned* __restrict__ dst) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int v[4] = {-1, -1, -1, -1}, cp[4];
    int cnt = (p < np) ? decode_pair(pairs[p], node_offset, v) : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= cnt || v[i] < 0 || v[i] >= n) cp[i] = -1;
        else cp[i] = (level < 0) ? real_map[v[i]] : coarse_space[(size_t)(level - 1) * n + v[i]];
    }
    unsigned msk[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i + 1; j < 4; ++j) {
            if (cp[i] < 0 || cp[j] < 0 || cp[i] == cp[j]) continue;
            if (cp[i] / B == cp[j] / B) {
                msk[i] |= 1u << (cp[j] % B);
                msk[j] |= 1u << (cp[i] % B);
            }
        }
    // Coarse levels funnel many pairs onto few nodes: merge equal targets
    // across the wave (a few leader rounds) and skip ORs that add no bit.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int addr = (cp[i] >= 0 && msk[i] != 0u) ? (level < 0 ? v[i] : cp[i]) : -1;
        unsigned m = msk[i];
        unsigned long long active = __ballot(addr >= 0);
        for (int round = 0; round < 4 && active; ++round) {
            int leader = __ffsll((long long)active) - 1;
            int a = __shfl(addr, leader);
            bool same = addr >= 0 && addr == a;
            unsigned mm = same ? m : 0u;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) mm |= __shfl_xor(mm, off);
            if ((int)(threadIdx.x & 63) == leader) {
                unsigned cur = __hip_atomic_load(dst + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((cur | mm) != cur) atomicOr(dst + a, mm);
            }
            active &= ~__ballot(same);
            if (same) addr = -1;
        }
        if (addr >= 0) {
            unsigned cur = __hip_atomic_load(dst + addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((cur | m) != cur) atomicOr(dst + addr, m);
        }
    }
}

// closure inside each level-0 bank, component count per bank
__global__ void __launch_bounds__(kTB) k_l0_closure(int M, int B, const int* __restrict__ part_map,
                                                     unsigned* __restrict__ fine_mask, int* __restrict__ prefix0) {
    __shared__ unsigned sm[kTB];
    __shared__ unsigned sc[kTB];
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int lane = threadIdx.x % B, bbase = threadIdx.x - lane;
    int v = (s < M) ? part_map[s] : -1;
    sm[threadIdx.x] = (v >= 0) ? fine_mask[v] : 0u;
    __syncthreads();
    unsigned cl = 0u;
    if (v >= 0) {
        cl = closure(sm[threadIdx.x], lane, sm + bbase);
        fine_mask[v] = cl;
    }
    sc[threadIdx.x] = (v >= 0 && (cl & ltmask(lane)) == 0u) ? 1u : 0u;
    __syncthreads();
    if (lane == 0 && s < M) {
        int cnt = 0;
        for (int l = 0; l < B; ++l) cnt += (int)sc[bbase + l];
        prefix0[s / B] = cnt;
    }
}

__global__ void __launch_bounds__(kTB) k_level1(int M, int Mpad, int n, int B, const int* __restrict__ part_map,
                                                 const unsigned* __restrict__ fine_mask,
                                                 const int* __restrict__ psum0, int* __restrict__ coarse_space,
                                                 int* __restrict__ going_next) {
    __shared__ unsigned sel[kTB];
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int lane = threadIdx.x % B, bbase = threadIdx.x - lane;
    int v = (s < M) ? part_map[s] : -1;
    unsigned fm = (v >= 0) ? fine_mask[v] : 0u;
    sel[threadIdx.x] = (v >= 0 && (fm & ltmask(lane)) == 0u) ? (1u << lane) : 0u;
    __syncthreads();
    if (v < 0) return;
    unsigned elected = 0u;
    for (int l = 0; l < B; ++l) elected |= sel[bbase + l];
    int el = lowbit(fm);
    int pre = __popc(elected & ltmask(el)) + psum0[s / B];
    coarse_space[v] = pre;
    going_next[v] = pre + Mpad;
}

__global__ void k_set_level_size(int* __restrict__ ls, int level, const int* __restrict__ total, int B,
                                 int first_begin) {
    // ls[2*(level+1)] = total comps, ls[2*(level+1)+1] = begin_level + pad(count_level)
    int cnt_prev = ls[2 * level + 0];
    int beg_prev = ls[2 * level + 1];
    ls[2 * (level + 1) + 0] = *total;
    ls[2 * (level + 1) + 1] = (level == 0) ? first_begin : beg_prev + (cnt_prev + B - 1) / B * B;
}

// coarse connectivity at `level` (nodes are level-local ids in coarse_space
// row level-1).  Masks are OR-reduced across the wavefront when every lane
// targets the same coarse node (the common case on coarse levels, where many
// vertices share one node), and an atomic is skipped when it would add no bit.
__global__ void k_lx_masks(int n, int B, int level, const unsigned* __restrict__ nstart,
                           const unsigned* __restrict__ nnum, const unsigned* __restrict__ nlist,
                           const int* __restrict__ coarse_space, unsigned* __restrict__ next_mask) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = v < n;
    const int* cs = coarse_space + (size_t)(level - 1) * n;
    int ci = valid ? cs[v] : -1;
    unsigned msk = 0u;
    if (valid) {
        unsigned st = nstart[v], nn = nnum[v];
        for (unsigned i0 = 0; i0 < nn; i0 += kNbrBatch) {  // as in k_l0_masks
            int ids[kNbrBatch], cu[kNbrBatch];
#pragma unroll
            for (int k = 0; k < kNbrBatch; ++k) ids[k] = (i0 + k < nn) ? (int)nlist[st + i0 + k] : -1;
#pragma unroll
            for (int k = 0; k < kNbrBatch; ++k) cu[k] = ids[k] >= 0 ? cs[ids[k]] : -1;
#pragma unroll
            for (int k = 0; k < kNbrBatch; ++k)
                if (ids[k] >= 0 && cu[k] / B == ci / B) msk |= 1u << (cu[k] % B);
        }
    }
    int first = __shfl(ci, __ffsll((long long)__ballot(valid)) - 1);
    bool uniform = __all(!valid || ci == first);
    if (uniform) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) msk |= __shfl_xor(msk, off);
        if ((threadIdx.x & 63) == __ffsll((long long)__ballot(valid)) - 1 && valid && msk) {
            unsigned cur = __hip_atomic_load(next_mask + ci, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((cur | msk) != cur) atomicOr(next_mask + ci, msk);
        }
    } else if (valid && msk) {
        unsigned cur = __hip_atomic_load(next_mask + ci, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((cur | msk) != cur) atomicOr(next_mask + ci, msk);
    }
}

__global__ void __launch_bounds__(kTB) k_lx_closure(int nmax, int B, int level, const int* __restrict__ ls,
                                                     unsigned* __restrict__ next_mask, int* __restrict__ next_prefix,
                                                     unsigned* __restrict__ comp_mask) {
    __shared__ unsigned sm[kTB];
    __shared__ unsigned sc[kTB];
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int lane = threadIdx.x % B, bbase = threadIdx.x - lane;
    int number = ls[2 * level + 0];
    bool valid = c < number;
    sm[threadIdx.x] = valid ? ((1u << lane) | next_mask[c]) : 0u;
    __syncthreads();
    unsigned cl = 0u;
    if (valid) cl = closure(sm[threadIdx.x], lane, sm + bbase);
    sc[threadIdx.x] = (valid && (cl & ltmask(lane)) == 0u) ? 1u : 0u;
    __syncthreads();
    if (valid) {
        next_mask[c] = cl;
        comp_mask[(size_t)ls[2 * level + 1] + c] = cl;  // component lanes of global node beg + c
    }
    if (lane == 0 && c < nmax) {
        int cnt = 0;
        for (int l = 0; l < B; ++l) cnt += (int)sc[bbase + l];
        next_prefix[c / B] = cnt;
    }
}

__global__ void __launch_bounds__(kTB) k_lx_renumber(int B, int level, const int* __restrict__ ls,
                                                      unsigned* __restrict__ next_mask,
                                                      const int* __restrict__ next_psum, int* __restrict__ going_next,
                                                      size_t going_cap) {
    __shared__ unsigned sel[kTB];
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int lane = threadIdx.x % B, bbase = threadIdx.x - lane;
    int number = ls[2 * level + 0], begin = ls[2 * level + 1];
    int pad = (number + B - 1) / B * B;
    bool valid = c < number;
    unsigned cm = valid ? next_mask[c] : 0u;
    sel[threadIdx.x] = (valid && (cm & ltmask(lane)) == 0u) ? (1u << lane) : 0u;
    __syncthreads();
    if (!valid) return;
    unsigned elected = 0u;
    for (int l = 0; l < B; ++l) elected |= sel[bbase + l];
    int pre = __popc(elected & ltmask(lowbit(cm))) + next_psum[c / B];
    next_mask[c] = (unsigned)pre;  // every lane of the block read its mask before the barrier
    size_t g = (size_t)begin + c;
    if (g < going_cap) going_next[g] = pre + begin + pad;
}

__global__ void k_next_table(int n, int level, int* __restrict__ coarse_space, const unsigned* __restrict__ next_mask) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int nx = coarse_space[(size_t)(level - 1) * n + v];
    coarse_space[(size_t)level * n + v] = (int)next_mask[nx];
}

__global__ void k_aggregate(int n, int L, const int* __restrict__ real_map, const int* __restrict__ going_next,
                            int* __restrict__ coarse_table, int* __restrict__ node_of) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int cur = v;
    node_of[v] = real_map[v];
    for (int l = 0; l < L - 1; ++l) {
        int nx = going_next[cur];
        coarse_table[6 * (size_t)v + l] = nx;
        node_of[(size_t)(l + 1) * n + v] = nx;
        cur = nx;
    }
    for (int l = L - 1; l < 6; ++l) coarse_table[6 * (size_t)v + l] = 0;
}

// ---- Galerkin projection ----------------------------------------------------
__device__ inline int pack_index(int B, int r, int c) { return B * r - r * (r + 1) / 2 + c; }  // r <= c

// Projection of one BCOO block H = A(r, c) (r <= c, column-major) onto the
// level-l bank entry of its node pair: the packed index and the 3x3 block
// (H + H^T on a coarse diagonal, H^T below the bank diagonal).
__device__ inline int project_block(int B, int li, int lj, bool same_vertex, const double* H, double* blk) {
    if (li == lj) {
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) blk[p + 3 * q] = same_vertex ? H[p + 3 * q] : H[p + 3 * q] + H[q + 3 * p];
        return pack_index(B, li, li);
    }
    if (li < lj) {
#pragma unroll
        for (int e = 0; e < 9; ++e) blk[e] = H[e];
        return pack_index(B, li, lj);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) blk[p + 3 * q] = H[q + 3 * p];
    return pack_index(B, lj, li);
}

// Galerkin projection for 16-vertex banks when vertex ids follow slot order:
// workgroup b owns the vertices whose level-1 node lies in level-1 bank b,
// i.e. the contiguous BCOO rows [vlo_b, vhi_b).  Level 0 targets are unique
// (plain stores).  Level-1 and level-2 targets accumulate in one-bank LDS
// windows (the first owned vertex's banks, which hold nearly every target)
// flushed once; the top levels (>= top_lvl, at most kProjTopBanksMax banks)
// accumulate in LDS and are written as per-workgroup partials, summed
// afterwards in a fixed order (k_project_top_reduce).  Other levels and
// window misses use global FP64 atomics.
// A block whose two vertices share the first owned vertex's level-2 node
// (most blocks: that node usually covers every owned vertex) adds the same
// 3x3 block, H + H^T (H on the diagonal), to that node's diagonal block on
// every level from 2 up, since nodes nest; such blocks are summed in
// registers and once over the workgroup instead of with LDS atomics that
// all lanes of a wavefront would aim at the same nine words.
constexpr int kProjWin1 = 1, kProjWin2 = 1;  // LDS window banks of levels 1 and 2
constexpr int kProjWinBanks = kProjWin1 + kProjWin2;
constexpr int kProjWinLevels = 2;
constexpr int kProjBank16 = 136 * 9;  // doubles per packed 16-vertex bank
constexpr int kProjTopBanks = 6;      // with the windows: 78 KB of LDS, two workgroups per CU
constexpr int kProjTopBanksMax = 12;  // larger top levels: 137 KB, one workgroup per CU
template <int TOPB>
__global__ void __launch_bounds__(512) k_project_banked(
    const int* __restrict__ ustart, const int* __restrict__ rows, const int* __restrict__ cols,
    const double* __restrict__ vals, int off, int n, int L, const int* __restrict__ node_of,
    const int* __restrict__ vlo, const int* __restrict__ vhi, int nbanks, double* __restrict__ bank_upper,
    int win_levels, int top_lvl, int top_lo, int top_nb, double* __restrict__ top_scratch) {
    constexpr int B = 16, NB = kProjBank16;
    __shared__ double win[kProjWinBanks * NB];  // level 1, then level 2
    __shared__ double top[TOPB * NB];
    __shared__ double red[9][512 / 64];
    __shared__ int wbase[kProjWinLevels];
    const int b = blockIdx.x, t = threadIdx.x;
    for (int i = t; i < kProjWinBanks * NB; i += blockDim.x) win[i] = 0.0;
    for (int i = t; i < top_nb * NB; i += blockDim.x) top[i] = 0.0;
    const int v0 = vlo[b], v1 = vhi[b];
    if (t < kProjWinLevels) wbase[t] = (v0 < v1 && t + 1 < L) ? node_of[(size_t)(t + 1) * n + v0] / B : -1;
    const int dom2 = (v0 < v1 && L > 2) ? node_of[2 * (size_t)n + v0] : -1;
    // LDS (or global) address of packed entry pidx of bank `bank` at level l >= 1
    auto target = [&](int l, int bank, int pidx) -> double* {
        if (l >= top_lvl) return top + (size_t)(bank - top_lo) * NB + 9 * pidx;
        if (l <= win_levels) {
            const int w = bank - wbase[l - 1];
            if (l == 1 && w == 0) return win + 9 * pidx;
            if (l == 2 && w >= 0 && w < kProjWin2) return win + (size_t)(kProjWin1 + w) * NB + 9 * pidx;
        }
        return bank_upper + (size_t)bank * NB + 9 * pidx;
    };
    double acc[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = 0.0;
    __syncthreads();
    if (v0 < v1) {
        const int k0 = ustart[v0 + off], k1 = ustart[v1 + off];
        for (int k = k0 + t; k < k1; k += blockDim.x) {
            const int r = rows[k] - off, c = cols[k] - off;
            if (c < 0 || c >= n) continue;
            double H[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) H[e] = vals[9 * (size_t)k + e];
            const bool dom = dom2 >= 0 && node_of[2 * (size_t)n + r] == dom2 && node_of[2 * (size_t)n + c] == dom2;
            const int lend = dom ? 2 : L;  // levels >= 2 of a dominant block: summed below
            for (int l = 0; l < lend; ++l) {
                const int I = node_of[(size_t)l * n + r], J = node_of[(size_t)l * n + c];
                if (I / B != J / B) continue;
                const int bank = I / B;
                double blk[9];
                const int pidx = project_block(B, I % B, J % B, r == c, H, blk);
                if (l == 0) {
                    double* dst = bank_upper + (size_t)bank * NB + 9 * pidx;  // unique per BCOO block
#pragma unroll
                    for (int e = 0; e < 9; ++e) dst[e] += blk[e];
                    continue;
                }
                double* dst = target(l, bank, pidx);
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    if (blk[e] != 0.0) atomicAdd(dst + e, blk[e]);
            }
            if (dom) {
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int q = 0; q < 3; ++q) acc[p + 3 * q] += (r == c) ? H[p + 3 * q] : H[p + 3 * q] + H[q + 3 * p];
            }
        }
    }
    // the dominant blocks' sum, over the workgroup, onto the diagonal block
    // of the first owned vertex's node on every level from 2 up
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        const double s = wave_sum(acc[e]);
        if ((t & 63) == 0) red[e][t >> 6] = s;
    }
    __syncthreads();
    if (t == 0 && dom2 >= 0) {
        double s[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            s[e] = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s[e] += red[e][w];
        }
        for (int l = 2; l < L; ++l) {
            const int I = node_of[(size_t)l * n + v0];
            double* dst = target(l, I / B, pack_index(B, I % B, I % B));
#pragma unroll
            for (int e = 0; e < 9; ++e)
                if (s[e] != 0.0) atomicAdd(dst + e, s[e]);
        }
    }
    __syncthreads();
    for (int q = t; q < kProjWinBanks * NB; q += blockDim.x) {
        const double v = win[q];
        const int w = q / NB, e = q % NB;
        const int lw = w < kProjWin1 ? 0 : 1;
        if (lw >= win_levels) continue;
        const int bank = wbase[lw] + (w < kProjWin1 ? w : w - kProjWin1);
        if (v != 0.0 && bank >= 0 && bank < nbanks) atomicAdd(bank_upper + (size_t)bank * NB + e, v);
    }
    double* out = top_scratch + (size_t)b * top_nb * NB;
    for (int q = t; q < top_nb * NB; q += blockDim.x) out[q] = top[q];
}

// Deterministic projection (opt-in): every BCOO block writes its level-0
// target directly and, for each coarse level where its node pair shares a
// bank, one (entry, block) record; a stable radix sort by entry groups the
// records in block order, and a fixed-order segmented sum (tiles of
// kProjTile records, see below) adds each group.  Same result on every run,
// for any bank size or numbering.
__global__ void k_proj_records(int K, const int* __restrict__ rows, const int* __restrict__ cols,
                               const double* __restrict__ vals, int off, int n, int B, int L,
                               const int* __restrict__ node_of, double* __restrict__ bank_upper, unsigned sentinel,
                               unsigned* __restrict__ key, unsigned* __restrict__ rec) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int npack = B * (B + 1) / 2;
    const int r = rows[k] - off, c = cols[k] - off;
    const bool in = r >= 0 && c >= 0 && r < n && c < n;
    for (int l = 1; l < L; ++l) {
        unsigned kk = sentinel;
        if (in) {
            const int I = node_of[(size_t)l * n + r], J = node_of[(size_t)l * n + c];
            if (I / B == J / B) {
                const int li = I % B, lj = J % B;
                kk = (unsigned)(I / B) * (unsigned)npack + (unsigned)(li <= lj ? pack_index(B, li, lj) : pack_index(B, lj, li));
            }
        }
        key[(size_t)k * (L - 1) + (l - 1)] = kk;
        rec[(size_t)k * (L - 1) + (l - 1)] = (unsigned)k * 8u + (unsigned)l;
    }
    if (!in) return;
    const int I = node_of[r], J = node_of[c];
    if (I / B != J / B) return;
    double H[9], blk[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) H[e] = vals[9 * (size_t)k + e];
    const int pidx = project_block(B, I % B, J % B, r == c, H, blk);
    double* dst = bank_upper + ((size_t)(I / B) * npack + pidx) * 9;  // level 0: unique per BCOO block
#pragma unroll
    for (int e = 0; e < 9; ++e) dst[e] += blk[e];
}

// contribution (9 doubles) of every sorted record
__global__ void k_proj_contrib(long long m, const unsigned* __restrict__ key, const unsigned* __restrict__ rec,
                               const int* __restrict__ rows, const int* __restrict__ cols,
                               const double* __restrict__ vals, int off, int n, int B, const int* __restrict__ node_of,
                               unsigned sentinel, double* __restrict__ contrib) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || key[i] == sentinel) return;
    const unsigned q = rec[i];
    const int k = (int)(q >> 3), l = (int)(q & 7u);
    const int r = rows[k] - off, c = cols[k] - off;
    const int I = node_of[(size_t)l * n + r], J = node_of[(size_t)l * n + c];
    double H[9], blk[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) H[e] = vals[9 * (size_t)k + e];
    (void)project_block(B, I % B, J % B, r == c, H, blk);
#pragma unroll
    for (int e = 0; e < 9; ++e) contrib[9 * (size_t)i + e] = blk[e];
}

// Fixed-order segmented sum: records are cut into tiles of kProjTile; the
// first record of each (entry, tile) piece sums the piece in record order,
// then the first record of each entry sums its pieces in tile order.
constexpr int kProjTile = 256;
__global__ void k_proj_pieces(long long m, const unsigned* __restrict__ key, unsigned sentinel,
                              double* __restrict__ contrib) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const unsigned kk = key[i];
    if (kk == sentinel || (i % kProjTile != 0 && key[i - 1] == kk)) return;
    const long long tend = min(m, (i / kProjTile + 1) * kProjTile);
    double acc[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = contrib[9 * (size_t)i + e];
    for (long long j = i + 1; j < tend && key[j] == kk; ++j)
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] += contrib[9 * (size_t)j + e];
#pragma unroll
    for (int e = 0; e < 9; ++e) contrib[9 * (size_t)i + e] = acc[e];  // the piece sum, at the piece head
}

__global__ void k_proj_entries(long long m, const unsigned* __restrict__ key, unsigned sentinel,
                               const double* __restrict__ contrib, double* __restrict__ bank_upper) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const unsigned kk = key[i];
    if (kk == sentinel || (i > 0 && key[i - 1] == kk)) return;
    double acc[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = contrib[9 * (size_t)i + e];
    for (long long j = (i / kProjTile + 1) * kProjTile; j < m && key[j] == kk; j += kProjTile)
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] += contrib[9 * (size_t)j + e];
    double* dst = bank_upper + (size_t)kk * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) dst[e] = acc[e];
}

// top-level bank entries = sum of the per-workgroup partials, fixed order
// (four interleaved partial sums per entry, combined in order)
__global__ void __launch_bounds__(256) k_project_top_reduce(const double* __restrict__ scratch, int nwg, int len,
                                                            double* __restrict__ dst) {
    __shared__ double part[4][64];
    const int e = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
    double s = 0.0;
    if (e < len)
        for (int w = q; w < nwg; w += 4) s += scratch[(size_t)w * len + e];
    part[q][threadIdx.x & 63] = s;
    __syncthreads();
    if (q == 0 && e < len) dst[e] = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                                    part[3][threadIdx.x];
}

// vertex range owned by each level-1 bank (slot-ordered vertex ids)
__global__ void k_owned_vertices(int nb0, int B, const int* __restrict__ prefix0, const int* __restrict__ psum0,
                                 const int* __restrict__ part_map, int* __restrict__ vlo, int* __restrict__ vhi) {
    int cb = blockIdx.x * blockDim.x + threadIdx.x;
    if (cb >= nb0 || prefix0[cb] <= 0) return;
    int mn = 0x7fffffff, mx = -1;
    for (int i = 0; i < B; ++i) {
        int v = part_map[cb * B + i];
        if (v >= 0) { mn = min(mn, v); mx = max(mx, v); }
    }
    if (mx < 0) return;
    int f = psum0[cb] / B;
    atomicMin(vlo + f, mn);
    atomicMax(vhi + f, mx + 1);
}

// element (r, c), r <= c, of a symmetric D x D matrix stored as its upper
// triangle packed by rows
template <int D>
__device__ inline int sym_index(int r, int c) { return r * (2 * D - r + 1) / 2 + (c - r); }

// ---- bank inverse: unpivoted in-place Gauss-Jordan in LDS -------------------
template <int B>
__global__ void __launch_bounds__(B == 16 ? 64 : 128) k_bank_inverse(int nbanks, const double* __restrict__ bank_upper,
                                                                      double* __restrict__ inv64,
                                                                      float* __restrict__ inv32) {
    constexpr int D = 3 * B;
    constexpr int npack = B * (B + 1) / 2;
    __shared__ double S[D][D + 1];
    __shared__ double colm[D];
    __shared__ double rt_s;
    int b = blockIdx.x;
    int i = threadIdx.x;  // column owned by this thread
    const double* src = bank_upper + (size_t)b * npack * 9;
    if (i < D) {
        int ci = i / 3, qi = i % 3;
        for (int j = 0; j < D; ++j) {
            int rj = j / 3, pj = j % 3;
            double v;
            if (ci >= rj) v = src[(size_t)pack_index(B, rj, ci) * 9 + pj + 3 * qi];   // block (rj,ci) elem (pj,qi)
            else v = src[(size_t)pack_index(B, ci, rj) * 9 + qi + 3 * pj];            // transpose of block (ci,rj)
            if (i == j && v == 0.0) v = 1.0;
            S[j][i] = v;
        }
    }
    __syncthreads();
    for (int j = 0; j < D; ++j) {
        if (i == 0) rt_s = S[j][j];
        if (i < D) colm[i] = S[i][j];
        __syncthreads();
        double rt = rt_s;
        if (i < D) S[i][j] = (i == j) ? 1.0 : 0.0;
        __syncthreads();
        if (i < D) {
            double pj = S[j][i] / rt;
            S[j][i] = pj;
            for (int k = 0; k < D; ++k)
                if (k != j) S[k][i] += -colm[k] * pj;
        }
        __syncthreads();
    }
    // symmetric result: keep the upper triangle, packed by rows
    if (i < D) {
        constexpr int NP = D * (D + 1) / 2;
        double* o64 = inv64 + (size_t)b * NP;
        float* o32 = inv32 + (size_t)b * NP;
        for (int r = 0; r <= i; ++r) {
            int e = sym_index<D>(r, i);
            o64[e] = S[r][i];
            o32[e] = (float)S[r][i];
        }
    }
}

// Same elimination for 16-vertex banks (D = 48) with the matrix held in
// registers: thread i owns column i; at step j the owner of column j
// publishes it (double-buffered LDS, one barrier per step) and every thread
// updates its own column.  Steps are instantiated per compile-time j so every
// register index is static.  Arithmetic and order equal k_bank_inverse's.
template <int J, int D>
__device__ __forceinline__ void gj_step(double (&S)[D], double (*colm)[D], int i) {
    double* cm = colm[J & 1];
    if (i == J) {
#pragma unroll
        for (int k = 0; k < D; ++k) cm[k] = S[k];
    }
    __syncthreads();
    const double rt = cm[J];
    if (i == J) {
#pragma unroll
        for (int k = 0; k < D; ++k) S[k] = (k == J) ? 1.0 : 0.0;
    }
    const double pv = S[J] / rt;
    S[J] = pv;
#pragma unroll
    for (int k = 0; k < D; ++k)
        if (k != J) S[k] += -cm[k] * pv;
}

template <int D, int... J>
__device__ __forceinline__ void gj_all(double (&S)[D], double (*colm)[D], int i, std::integer_sequence<int, J...>) {
    (gj_step<J, D>(S, colm, i), ...);
}

// value v of lane `lane` (a compile-time constant) in every lane, through
// scalar registers
template <int LANE>
__device__ __forceinline__ double lane_value(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffu), LANE);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), LANE);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The same step with the pivot column broadcast lane to lane (no LDS, no
// barrier).  Each entry k of the owner's column is read before the owner
// overwrites it in the same iteration; the owner starts from the unit
// vector e_J, as in gj_step, so the arithmetic is identical.
// (The owner's base is its entry times 0.0 -- one multiply instead of a
// 64-bit select; the same value up to the sign of a zero.)
template <int J, int D, int K>
__device__ __forceinline__ void gj_rl_entry(double (&S)[D], double pv, double keep) {
    if constexpr (K != J) {
        const double c = lane_value<J>(S[K]);
        const double base = S[K] * keep;
        S[K] = base + -c * pv;
    }
}
template <int J, int D, int... K>
__device__ __forceinline__ void gj_rl_step(double (&S)[D], int i, std::integer_sequence<int, K...>) {
    const bool own = i == J;
    const double keep = own ? 0.0 : 1.0;
    const double rt = lane_value<J>(S[J]);
    const double pv = (own ? 1.0 : S[J]) / rt;
    (gj_rl_entry<J, D, K>(S, pv, keep), ...);
    S[J] = pv;
}
template <int D, int... J>
__device__ __forceinline__ void gj_rl_all(double (&S)[D], int i, std::integer_sequence<int, J...>) {
    (gj_rl_step<J, D>(S, i, std::make_integer_sequence<int, D>{}), ...);
}

constexpr bool kBankInverseReadlane = true;  // pivot column by lane broadcast (false: through LDS)
__global__ void __launch_bounds__(64) k_bank_inverse_reg48(int nbanks, const double* __restrict__ bank_upper,
                                                           double* __restrict__ inv64, float* __restrict__ inv32) {
    constexpr int B = 16, D = 48, npack = B * (B + 1) / 2, NP = D * (D + 1) / 2;
    __shared__ double colm[2][D];
    const int b = blockIdx.x, i = threadIdx.x;
    const bool act = i < D;
    const double* src = bank_upper + (size_t)b * npack * 9;
    double S[D];
    const int ci = i / 3, qi = i % 3;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const int rj = j / 3, pj = j % 3;
        double v = 0.0;
        if (act) {
            v = (ci >= rj) ? src[(size_t)pack_index(B, rj, ci) * 9 + pj + 3 * qi]
                           : src[(size_t)pack_index(B, ci, rj) * 9 + qi + 3 * pj];
            if (i == j && v == 0.0) v = 1.0;
        }
        S[j] = v;
    }
    if (kBankInverseReadlane) {
        (void)colm;
        gj_rl_all<D>(S, i, std::make_integer_sequence<int, D>{});
    } else {
        gj_all<D>(S, colm, i, std::make_integer_sequence<int, D>{});
    }
    if (act) {
        double* o64 = inv64 + (size_t)b * NP;
        float* o32 = inv32 + (size_t)b * NP;
#pragma unroll
        for (int r = 0; r < D; ++r)
            if (r <= i) {
                const int e = sym_index<D>(r, i);
                o64[e] = S[r];
                o32[e] = (float)S[r];
            }
    }
}

// Half-storage form of the packed 48 x 48 inverses for the fused FP32 bank
// solves: D = sqrt of the inverse's diagonal (1 where it is not positive),
// S = D^-1 inv D^-1 rounded to FP16 (unit diagonal; |S_ij| <= 1 for an SPD
// inverse, clamped to the FP16 range otherwise).  One wavefront per bank,
// FP64 arithmetic from the FP64 inverses when they were built.
template <typename S>
__global__ void __launch_bounds__(64) k_pack_inv_half(int nbanks, const S* __restrict__ inv,
                                                      unsigned short* __restrict__ h, float* __restrict__ dsc) {
    constexpr int D = 48, NP = D * (D + 1) / 2;
    __shared__ double d[D];
    const int b = blockIdx.x, i = threadIdx.x;
    if (b >= nbanks) return;
    const S* src = inv + (size_t)b * NP;
    if (i < D) {
        const double a = (double)src[pack_index(D, i, i)];
        const double di = a > 0.0 ? sqrt(a) : 1.0;
        d[i] = di;
        dsc[(size_t)b * D + i] = (float)di;
    }
    __syncthreads();
    unsigned short* dst = h + (size_t)b * NP;
    for (int r = 0; r < D; ++r) {
        const int c = r + i;
        if (c < D) {
            const int q = pack_index(D, r, c);
            const float v = (float)((double)src[q] / (d[r] * d[c]));
            const _Float16 hv = (_Float16)fminf(fmaxf(v, -65504.f), 65504.f);
            dst[q] = __builtin_bit_cast(unsigned short, hv);
        }
    }
}

// ---- apply -----------------------------------------------------------------
// value of v in lane j of the wavefront (j uniform)
__device__ inline float lane_bcast(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ inline double lane_bcast(double v, int j) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// address-space-qualified pointers for the direct global -> LDS load builtin
typedef __attribute__((address_space(1))) const void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;

// z_bank = A^-1 r_bank for every bank of every level: the packed inverse is
// staged through LDS (the bytes of this kernel are almost all inverse
// bytes), then lane i forms row i, summing over columns in order.
template <int B, typename T>
__global__ void __launch_bounds__(B == 16 ? 64 : 128) k_bank_solve(const T* __restrict__ inv, const T* __restrict__ R,
                                                                    T* __restrict__ Z, const int* __restrict__ ctrl,
                                                                    SceneSkip skip) {
    if (ctrl && ctrl[0]) return;
    if (skip.stopped(blockIdx.x)) return;
    constexpr int D = 3 * B;
    constexpr int NP = D * (D + 1) / 2;
    constexpr int NT = (B == 16) ? 64 : 128;
    constexpr int NV = NP * (int)sizeof(T) / 16;  // exact: NP * sizeof(T) is a multiple of 16 for B = 16, 32
    static_assert(NP * sizeof(T) % 16 == 0, "packed bank must be a whole number of 16-byte vectors");
    using V4 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    constexpr int PER = (NV + NT - 1) / NT;
    __shared__ V4 As[PER * NT];  // padded to whole 1 KB wave chunks
    __shared__ T rs[B == 16 ? 1 : D];
    const int b = blockIdx.x, i = threadIdx.x;
    const V4* src = reinterpret_cast<const V4*>(inv + (size_t)b * NP);
    // Direct global -> LDS loads (16 B per lane, one 1 KB chunk per wave and
    // instruction), all in flight together; lanes past the bank re-read its
    // last vector into the padding.
    const int wv = __builtin_amdgcn_readfirstlane(i / 64), lane = i % 64;
    T rv = R[(size_t)b * D + min(i, D - 1)];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q0 = k * NT + wv * 64;
        __builtin_amdgcn_global_load_lds((gptr_t)(src + min(q0 + lane, NV - 1)), (lptr_t)(As + q0), 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    if constexpr (B != 16) {
        if (i < D) rs[i] = rv;
    }
    __syncthreads();
    if (i >= D) return;
    const T* A = reinterpret_cast<const T*>(As);
    // Row i of the symmetric bank from the packed upper triangle: element
    // (j, i) for j <= i, then (i, j); the packed index advances by D - 1 - j
    // while j < i and by 1 after.  r_j comes from lane j's register when the
    // bank fits one wavefront (B = 16), from LDS otherwise.
    T acc = 0;
    int idx = i;  // (0, i)
#pragma unroll 8
    for (int j = 0; j < D; ++j) {
        T rj;
        if constexpr (B == 16) rj = lane_bcast(rv, j);
        else rj = rs[j];
        acc += A[idx] * rj;
        idx += (j < i) ? (D - 1 - j) : 1;
    }
    Z[(size_t)b * D + i] = acc;
}

template <typename T>
__global__ void k_collect(int n, int L, const int* __restrict__ node_of, const T* __restrict__ Z,
                          double* __restrict__ z) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    T s0 = 0, s1 = 0, s2 = 0;
    for (int l = 0; l < L; ++l) {
        size_t g = (size_t)node_of[(size_t)l * n + v];
        s0 += Z[3 * g]; s1 += Z[3 * g + 1]; s2 += Z[3 * g + 2];
    }
    z[3 * (size_t)v] = (double)s0; z[3 * (size_t)v + 1] = (double)s1; z[3 * (size_t)v + 2] = (double)s2;
}

// Component sum of one bank-aligned group of nodes held in LDS (one node
// per thread).  `mask` holds the bank lanes of this node's component (its
// closure, self included); only the lowest lane of the component sums, in
// ascending lane order.  Returns true for that lane, with the sums.
template <typename T>
__device__ inline bool component_sum(unsigned mask, const T* a0, const T* a1, const T* a2, int B, T& s0, T& s1,
                                     T& s2) {
    int t = threadIdx.x, lane = t % B, bbase = t - lane;
    if (mask & ltmask(lane)) return false;
    s0 = 0; s1 = 0; s2 = 0;
    for (unsigned m = mask; m; m &= m - 1u) {
        int l = lowbit(m);
        s0 += a0[bbase + l]; s1 += a1[bbase + l]; s2 += a2[bbase + l];
    }
    return true;
}

// level-0 slots -> level-1 sums of the nodes of level-1 bank b, then
// level 1 -> 2 inside that bank.  `value(v, sl, own, ...)` gives the
// level-0 value of vertex v at slot sl.  A vertex is owned by exactly one
// workgroup: the one whose level-1 bank holds the vertex's level-1 node.
// Only the owner reads or writes the vertex's data (value() is called with
// own = false for the others and their values are never summed: a
// component's lanes all share one level-1 node), so a level-0 bank whose
// coarse nodes straddle two level-1 banks is split between their
// workgroups with no slot touched twice.  fine_mask / comp_mask: component
// lanes of each level-0 vertex / coarse node.
// The index loads of level-1 bank b that do not depend on the residual:
// its level-0 slot range, the first slot chunk's vertices and the level
// 1 -> 2 mapping of its nodes.  Issued before a kernel's partial-sum
// reduction, their latency overlaps it.
struct BankIdx {
    int lo, hi, v0, q;
    unsigned mask2;
    bool valid;
};
__device__ inline BankIdx bank_index(int b, int B, int L, int cnt1, int beg1, const int* __restrict__ part_map,
                                     const int* __restrict__ going_next, const unsigned* __restrict__ comp_mask,
                                     const int* __restrict__ l1lo, const int* __restrict__ l1hi) {
    const int t = threadIdx.x;
    BankIdx x;
    const int loc2 = b * B + t;
    x.valid = L >= 3 && t < B && loc2 < cnt1;
    x.q = -1;
    x.mask2 = 0u;
    if (x.valid) {
        x.q = going_next[(size_t)beg1 + loc2];
        x.mask2 = comp_mask[(size_t)beg1 + loc2];
    }
    x.lo = l1lo[b] * B;
    x.hi = l1hi[b] * B;
    const int sl = x.lo + t;
    x.v0 = (sl < x.hi) ? part_map[sl] : -1;
    return x;
}

template <typename T, typename F>
__device__ inline void restrict_bank(int b, int B, int L, int cnt1, int beg1, const int* __restrict__ part_map,
                                     const int* __restrict__ going_next, const unsigned* __restrict__ fine_mask,
                                     const unsigned* __restrict__ comp_mask, const BankIdx& pre, T* __restrict__ R,
                                     T (*sv)[kTB], T (*l1v)[32], F value) {
    const int t = threadIdx.x;
    const bool valid = pre.valid;
    const int q = pre.q;
    const unsigned mask2 = pre.mask2;
    const int lo = pre.lo, hi = pre.hi;
    for (int c0 = lo; c0 < hi; c0 += kTB) {
        int sl = c0 + t;
        int v = c0 == lo ? pre.v0 : ((sl < hi) ? part_map[sl] : -1);
        int p = -1;
        unsigned mask = 0u;
        T v0 = 0, v1 = 0, v2 = 0;
        if (v >= 0) {
            const int loc = going_next[v] - beg1;
            const bool own = loc / B == b;
            value(v, sl, own, v0, v1, v2);
            if (own) {
                R[3 * (size_t)sl] = v0; R[3 * (size_t)sl + 1] = v1; R[3 * (size_t)sl + 2] = v2;
                p = loc;
                mask = fine_mask[v];
            }
        }
        sv[0][t] = v0; sv[1][t] = v1; sv[2][t] = v2;
        __syncthreads();
        T s0, s1, s2;
        if (p >= 0 && component_sum<T>(mask, sv[0], sv[1], sv[2], B, s0, s1, s2)) {
            size_t g = (size_t)beg1 + p;
            R[3 * g] = s0; R[3 * g + 1] = s1; R[3 * g + 2] = s2;
            l1v[0][p - b * B] = s0; l1v[1][p - b * B] = s1; l1v[2][p - b * B] = s2;
        }
        __syncthreads();
    }
    if (L < 3) return;
    if (t < B) {
        sv[0][t] = valid ? l1v[0][t] : (T)0;
        sv[1][t] = valid ? l1v[1][t] : (T)0;
        sv[2][t] = valid ? l1v[2][t] : (T)0;
    }
    __syncthreads();
    T s0, s1, s2;
    if (valid && component_sum<T>(mask2, sv[0], sv[1], sv[2], B, s0, s1, s2)) {
        R[3 * (size_t)q] = s0; R[3 * (size_t)q + 1] = s1; R[3 * (size_t)q + 2] = s2;
    }
    __syncthreads();
}

// Multilevel restriction R_l = sum of children, levels 0 to 2 in one launch.
// Workgroup b owns level-1 bank b: it gathers the level-0 slots of every
// level-0 bank holding one of its children (reading the residual r through
// part_map, writing R_0 for the banks whose first child lands in b), forms
// its level-1 sums, and from them the level-2 nodes (whose children always
// lie in one level-1 bank).  Sum orders are those of the level-by-level
// component sums, so results do not depend on the schedule.
template <typename V, typename T>
__global__ void __launch_bounds__(kTB) k_restrict_all(const V* __restrict__ r, long long sub3, int B, int L,
                                                       const int* __restrict__ ls, const int* __restrict__ part_map,
                                                       const int* __restrict__ going_next,
                                                       const unsigned* __restrict__ fine_mask,
                                                       const unsigned* __restrict__ comp_mask,
                                                       const int* __restrict__ l1lo, const int* __restrict__ l1hi,
                                                       T* __restrict__ R,
                                                       const int* __restrict__ ctrl, SceneSkip skip) {
    if (ctrl && ctrl[0]) return;
    if (skip.bank_scene && skip.stopped(ls[3] / B + (int)blockIdx.x)) return;
    __shared__ T sv[3][kTB];
    __shared__ T l1v[3][32];
    auto value = [&](int v, int, bool own, T& v0, T& v1, T& v2) {
        if (!own) return;
        const V* rv = r + sub3 + 3 * (size_t)v;
        v0 = (T)rv[0]; v1 = (T)rv[1]; v2 = (T)rv[2];
    };
    const BankIdx pre = bank_index(blockIdx.x, B, L, ls[2], ls[3], part_map, going_next, comp_mask, l1lo, l1hi);
    restrict_bank<T>(blockIdx.x, B, L, ls[2], ls[3], part_map, going_next, fine_mask, comp_mask, pre, R, sv, l1v,
                     value);
}

// One PCG step fused with the restriction of its new residual:
// alpha = rz / p.Ap, x += alpha p, r -= alpha Ap, partials of r.r, and
// R_0..R_2 of the new r.  Workgroups [0, g1) walk the level-1 banks: each
// recomputes r for every vertex of its banks (identical arithmetic wherever
// a level-0 bank straddles two level-1 banks) and stores x, r and R_0 only
// for the banks it owns; workgroups [g1, grid) update the dofs outside the
// preconditioned range.  The iteration counter advances at the end.
template <typename V, typename T>
__global__ void __launch_bounds__(kTB) k_update_restrict(
    long long N, double* __restrict__ x, V* __restrict__ r, const V* __restrict__ p, const V* __restrict__ Ap,
    const double* __restrict__ part_pap, int np_pap, const double* __restrict__ part_rz, int np_rz,
    double* __restrict__ part_rr, int* __restrict__ ctrl, int sub_off, int n_mas, int B, int L,
    const int* __restrict__ ls, const int* __restrict__ part_map, const int* __restrict__ going_next,
    const unsigned* __restrict__ fine_mask, const unsigned* __restrict__ comp_mask, const int* __restrict__ l1lo,
    const int* __restrict__ l1hi, int nb1, int g1, T* __restrict__ R) {
    if (ctrl[C_STOP]) return;
    __shared__ double sh[2 * (kTB / 64)];
    __shared__ T sv[3][kTB];
    __shared__ T l1v[3][32];
    // this workgroup's first bank indices, loaded while the dots are reduced
    BankIdx pre{0, 0, -1, -1, 0u, false};
    if ((int)blockIdx.x < g1 && (int)blockIdx.x < nb1)
        pre = bank_index(blockIdx.x, B, L, ls[2], ls[3], part_map, going_next, comp_mask, l1lo, l1hi);
    double rz, pap;
    sum_partials2(part_rz, np_rz, part_pap, np_pap, sh, rz, pap);
    const double al = rz / pap;
    double acc = 0.0;
    if ((int)blockIdx.x < g1) {
        const int cnt1 = ls[2], beg1 = ls[3];
        // only the vertex's owner reads and updates r in place (see restrict_bank)
        auto value = [&](int v, int, bool own, T& v0, T& v1, T& v2) {
            if (!own) return;
            const size_t o = 3 * ((size_t)sub_off + v);
            V r0 = (V)((double)r[o] - al * (double)Ap[o]);
            V r1 = (V)((double)r[o + 1] - al * (double)Ap[o + 1]);
            V r2 = (V)((double)r[o + 2] - al * (double)Ap[o + 2]);
            x[o] += al * (double)p[o]; x[o + 1] += al * (double)p[o + 1]; x[o + 2] += al * (double)p[o + 2];
            r[o] = r0; r[o + 1] = r1; r[o + 2] = r2;
            acc += (double)r0 * (double)r0 + (double)r1 * (double)r1 + (double)r2 * (double)r2;
            v0 = (T)r0; v1 = (T)r1; v2 = (T)r2;
        };
        for (int b = blockIdx.x; b < nb1; b += g1)
            restrict_bank<T>(b, B, L, cnt1, beg1, part_map, going_next, fine_mask, comp_mask,
                             b == (int)blockIdx.x ? pre
                                                  : bank_index(b, B, L, cnt1, beg1, part_map, going_next, comp_mask,
                                                               l1lo, l1hi),
                             R, sv, l1v, value);
    } else {
        const long long m0 = 3LL * sub_off, m1 = 3LL * ((long long)sub_off + n_mas), nout = N - (m1 - m0);
        for (long long q = (long long)(blockIdx.x - g1) * kTB + threadIdx.x; q < nout;
             q += (long long)(gridDim.x - g1) * kTB) {
            long long i = q < m0 ? q : q + (m1 - m0);
            x[i] += al * (double)p[i];
            V ri = (V)((double)r[i] - al * (double)Ap[i]);
            r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    write_partial(acc, part_rr, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl[C_ITER] += 1;
}

// Levels 2 -> 3 -> ... -> L-1 (few nodes): one workgroup, level by level.
// A level's sums stay in LDS as the next level's input when that level has
// at most kTop nodes; otherwise they are read back from R (written by this
// workgroup before a barrier).
constexpr int kTop = 1024;
template <typename T>
__global__ void __launch_bounds__(kTop) k_restrict_top(int B, int L, const int* __restrict__ ls,
                                                        const int* __restrict__ going_next,
                                                        const unsigned* __restrict__ comp_mask, T* __restrict__ R,
                                                        const int* __restrict__ ctrl) {
    if (ctrl && ctrl[0]) return;
    __shared__ T sv[3][kTop];
    __shared__ T nx[3][kTop];
    const int t = threadIdx.x;
    bool in_lds = false;  // this level's values are in nx
    for (int l = 2; l + 1 < L; ++l) {
        const int cnt = ls[2 * l], beg = ls[2 * l + 1], cntn = ls[2 * l + 2], begn = ls[2 * l + 3];
        const bool next_fits = cntn <= kTop;
        for (int c0 = 0; c0 < cnt; c0 += kTop) {
            int cc = c0 + t;
            bool valid = cc < cnt;
            size_t g = (size_t)beg + cc;
            int pp = valid ? going_next[g] : -1;
            unsigned mask = valid ? comp_mask[g] : 0u;
            T v0 = 0, v1 = 0, v2 = 0;
            if (valid) {
                if (in_lds) { v0 = nx[0][cc]; v1 = nx[1][cc]; v2 = nx[2][cc]; }
                else { v0 = R[3 * g]; v1 = R[3 * g + 1]; v2 = R[3 * g + 2]; }
            }
            __syncthreads();  // nx fully read before this pass overwrites it
            sv[0][t] = v0; sv[1][t] = v1; sv[2][t] = v2;
            __syncthreads();
            T s0, s1, s2;
            if (valid && component_sum<T>(mask, sv[0], sv[1], sv[2], B, s0, s1, s2)) {
                size_t q = (size_t)pp;
                R[3 * q] = s0; R[3 * q + 1] = s1; R[3 * q + 2] = s2;
                if (next_fits) { int j = (int)(q - begn); nx[0][j] = s0; nx[1][j] = s1; nx[2][j] = s2; }
            }
            __syncthreads();
        }
        in_lds = next_fits;
    }
}

// ---- top-level restriction overlapped with the bank solves (16-vertex banks) -
// k_mas_solve_top replaces k_restrict_top + k_bank_solve in a PCG iteration
// (the x/r update before it restricts levels 0..2; k_collect_dot follows):
// workgroup 0 restricts levels 2 -> L-1 (the sums of k_restrict_top) and
// then solves the banks of levels >= 3 on its own results; the other
// workgroups solve every bank of levels 0..2, whose restriction is complete
// at launch, at the same time.  No workgroup waits for another.
constexpr int kCW = 8;               // wavefronts per workgroup of k_mas_coarse
constexpr int kCT = 64 * kCW;

// One 48 x 48 bank solve by one wavefront: the packed FP32/FP64 inverse is
// staged into this wavefront's LDS area by direct global -> LDS loads, lane
// i < 48 forms row i (the arithmetic of k_bank_solve<16>).
template <typename T>
__device__ inline T wave_bank_solve(const T* __restrict__ inv, const T* __restrict__ R, int b, void* lds) {
    constexpr int D = 48, NP = D * (D + 1) / 2;
    using V4 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    constexpr int NV = NP * (int)sizeof(T) / 16;
    constexpr int PER = (NV + 63) / 64;
    V4* As = static_cast<V4*>(lds);
    const int i = threadIdx.x & 63;
    const V4* src = reinterpret_cast<const V4*>(inv + (size_t)b * NP);
    const T rv = R[(size_t)b * D + min(i, D - 1)];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q0 = k * 64;
        __builtin_amdgcn_global_load_lds((gptr_t)(src + min(q0 + i, NV - 1)), (lptr_t)(As + q0), 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const T* A = reinterpret_cast<const T*>(As);
    T acc = 0;
    if (i < D) {
        int idx = i;
#pragma unroll 8
        for (int j = 0; j < D; ++j) {
            acc += A[idx] * lane_bcast(rv, j);
            idx += (j < i) ? (D - 1 - j) : 1;
        }
    }
    __builtin_amdgcn_wave_barrier();  // every lane's reads of this bank done before the area is reused
    return acc;
}

// The same solve from the half-storage form of an inverse: inv = D S D with
// D = diag(sqrt(inv_ii)) (FP32, one per bank row) and S (unit diagonal,
// |S_ij| <= 1 for an SPD inverse) packed like inv as FP16, so the bank reads
// 2352 + 192 bytes instead of 4704.  z_i = d_i sum_j S_ij (d_j r_j), FP32
// arithmetic.
__device__ inline float wave_bank_solve_h(const unsigned short* __restrict__ inv, const float* __restrict__ dsc,
                                          const float* __restrict__ R, int b, void* lds) {
    constexpr int D = 48, NP = D * (D + 1) / 2;
    constexpr int NV = NP * 2 / 16;  // 147 16-byte pieces per bank (2352 bytes)
    constexpr int PER = (NV + 63) / 64;
    float4* As = static_cast<float4*>(lds);
    const int i = threadIdx.x & 63;
    const float4* src = reinterpret_cast<const float4*>(inv + (size_t)b * NP);
    const float di = dsc[(size_t)b * D + min(i, D - 1)];
    const float rv = R[(size_t)b * D + min(i, D - 1)] * di;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q0 = k * 64;
        __builtin_amdgcn_global_load_lds((gptr_t)(src + min(q0 + i, NV - 1)), (lptr_t)(As + q0), 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const _Float16* A = reinterpret_cast<const _Float16*>(As);
    float acc = 0.f;
    if (i < D) {
        int idx = i;
#pragma unroll 8
        for (int j = 0; j < D; ++j) {
            acc += (float)A[idx] * lane_bcast(rv, j);
            idx += (j < i) ? (D - 1 - j) : 1;
        }
    }
    __builtin_amdgcn_wave_barrier();
    return acc * di;
}

// one bank solve from either storage (H: the half form; inv is then the FP16
// S and dsc its row scales)
template <typename T, bool H>
__device__ inline T bank_solve_any(const void* __restrict__ inv, const float* __restrict__ dsc,
                                   const T* __restrict__ R, int b, void* lds) {
    if constexpr (H) return wave_bank_solve_h(static_cast<const unsigned short*>(inv), dsc, R, b, lds);
    else return wave_bank_solve<T>(static_cast<const T*>(inv), R, b, lds);
}

template <typename T, bool H>
__global__ void __launch_bounds__(kCT) k_mas_solve_top(int L, const int* __restrict__ ls,
                                                     const int* __restrict__ going_next,
                                                     const unsigned* __restrict__ comp_mask, int top_bank0,
                                                     int nbanks, const void* __restrict__ inv,
                                                     const float* __restrict__ dsc, T* __restrict__ R,
                                                     T* __restrict__ Z, const int* __restrict__ ctrl, SceneSkip skip) {
    static_assert(!H || sizeof(T) == 4, "half-storage inverses serve FP32 applies");
    if (ctrl[C_STOP]) return;
    constexpr int D = 48;
    using V4 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    constexpr int PER = ((D * (D + 1) / 2) * (int)sizeof(T) / 16 + 63) / 64;
    static_assert(kCW * PER * 64 * sizeof(V4) >= 6 * kCT * sizeof(T), "LDS union too small");
    __shared__ V4 lds[kCW][PER * 64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (blockIdx.x == 0) {
        // levels 2 -> L-1, one level at a time (the sums of k_restrict_top)
        T(*sv)[kCT] = reinterpret_cast<T(*)[kCT]>(&lds[0][0]);
        T(*nx)[kCT] = reinterpret_cast<T(*)[kCT]>(reinterpret_cast<char*>(&lds[0][0]) + 3 * kCT * sizeof(T));
        const int t = threadIdx.x;
        bool in_lds = false;
        for (int l = 2; l + 1 < L; ++l) {
            const int cnt = ls[2 * l], beg = ls[2 * l + 1], cntn = ls[2 * l + 2], begn = ls[2 * l + 3];
            const bool next_fits = cntn <= kCT;
            for (int c0 = 0; c0 < cnt; c0 += kCT) {
                const int cc = c0 + t;
                const bool valid = cc < cnt;
                const size_t g = (size_t)beg + cc;
                const int pp = valid ? going_next[g] : -1;
                const unsigned mask = valid ? comp_mask[g] : 0u;
                T v0 = 0, v1 = 0, v2 = 0;
                if (valid) {
                    if (in_lds) { v0 = nx[0][cc]; v1 = nx[1][cc]; v2 = nx[2][cc]; }
                    else { v0 = R[3 * g]; v1 = R[3 * g + 1]; v2 = R[3 * g + 2]; }
                }
                __syncthreads();
                sv[0][t] = v0; sv[1][t] = v1; sv[2][t] = v2;
                __syncthreads();
                T s0, s1, s2;
                if (valid && component_sum<T>(mask, sv[0], sv[1], sv[2], 16, s0, s1, s2)) {
                    const size_t q = (size_t)pp;
                    R[3 * q] = s0; R[3 * q + 1] = s1; R[3 * q + 2] = s2;
                    if (next_fits) { const int j = (int)(q - begn); nx[0][j] = s0; nx[1][j] = s1; nx[2][j] = s2; }
                }
                __syncthreads();
            }
            in_lds = next_fits;
        }
        // the top banks, on this workgroup's restriction (R written above;
        // the barrier orders this workgroup's global writes and reads)
        __threadfence_block();
        __syncthreads();
        for (int b = top_bank0 + wv; b < nbanks; b += kCW) {
            if (skip.stopped(b)) continue;
            const T zb = bank_solve_any<T, H>(inv, dsc, R, b, &lds[wv][0]);
            if (lane < D) Z[(size_t)b * D + lane] = zb;
        }
        return;
    }
    const int b = (blockIdx.x - 1) * kCW + wv;  // levels 0 .. 2
    if (b < top_bank0 && !skip.stopped(b)) {
        const T zb = bank_solve_any<T, H>(inv, dsc, R, b, &lds[wv][0]);
        if (lane < D) Z[(size_t)b * D + lane] = zb;
    }
}

// level-0 bank range of each level-1 bank (children of its nodes)
__global__ void k_l1_ranges(int nb0, int B, const int* __restrict__ prefix0, const int* __restrict__ psum0,
                            int* __restrict__ lo, int* __restrict__ hi) {
    int cb = blockIdx.x * blockDim.x + threadIdx.x;
    if (cb >= nb0 || prefix0[cb] <= 0) return;
    int f = psum0[cb] / B, e = (psum0[cb] + prefix0[cb] - 1) / B;
    for (int b = f; b <= e; ++b) {
        atomicMin(lo + b, cb);
        atomicMax(hi + b, cb + 1);
    }
}

int num_levels(int nn, int B) {
    int nLevel = 1;
    int levelSz = (nn + B - 1) / B * B;
    while (levelSz > B) {
        levelSz /= B;
        nLevel++;
        levelSz = (levelSz + B - 1) / B * B;
    }
    nLevel = nLevel + 1;
    return nLevel > 6 ? 6 : nLevel;
}

template <typename V, typename T>
void mas_restrict_t(Ctx& c, const V* r, long long sub3, T* R, const int* ctrl, bool top = true,
                    SceneSkip skip = SceneSkip{}) {
    MasState& m = c.mas;
    if (m.nb1 <= 0) return;
    hipLaunchKernelGGL((k_restrict_all<V, T>), dim3(m.nb1), dim3(kTB), 0, c.stream, r, sub3, m.bank, m.L,
                       (const int*)m.level_size.p, (const int*)m.part_map.p, (const int*)m.going_next.p,
                       (const unsigned*)m.fine_mask.p, (const unsigned*)m.comp_mask.p,
                       (const int*)m.l1lo.p, (const int*)m.l1hi.p, R, ctrl, skip);
    SGIPC_LAUNCH_CHECK();
    if (top && m.L >= 4) {
        hipLaunchKernelGGL((k_restrict_top<T>), dim3(1), dim3(kTop), 0, c.stream, m.bank, m.L,
                           (const int*)m.level_size.p, (const int*)m.going_next.p, (const unsigned*)m.comp_mask.p, R, ctrl);
        SGIPC_LAUNCH_CHECK();
    }
}

template <typename T>
void mas_solve_t(Ctx& c, const T* inv, const T* R, T* Z, const int* ctrl, SceneSkip skip = SceneSkip{}) {
    MasState& m = c.mas;
    prof_begin(c, PROF_MAS_SOLVE);
    if (m.bank == 16)
        hipLaunchKernelGGL((k_bank_solve<16, T>), dim3(m.nbanks), dim3(64), 0, c.stream, inv, R, Z, ctrl, skip);
    else
        hipLaunchKernelGGL((k_bank_solve<32, T>), dim3(m.nbanks), dim3(128), 0, c.stream, inv, R, Z, ctrl, skip);
    SGIPC_LAUNCH_CHECK();
    prof_end(c, PROF_MAS_SOLVE);
}

template <typename T>
void mas_apply_t(Ctx& c, const double* r, double* z, T* R, T* Z, const T* inv) {
    MasState& m = c.mas;
    hipStream_t s = c.stream;
    if (m.n <= 0) return;
    mas_restrict_t<double, T>(c, r, 0, R, nullptr);
    mas_solve_t<T>(c, inv, R, Z, nullptr);
    hipLaunchKernelGGL(k_collect<T>, dim3(div_up(m.n, 256)), dim3(256), 0, s, m.n, m.L, m.node_of.p, Z, z);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace

void mas_set_topology(Ctx& c, int n, int node_offset, const unsigned* nstart, const unsigned* nnum,
                      const unsigned* nlist, int nlist_len, const int* part_map, int M, const int* real_map,
                      int bank) {
    if (bank != 16 && bank != 32) throw arg_error("MAS bank size must be 16 or 32");
    if (n < 0 || M < 0 || M % bank != 0) throw arg_error("MAS map size must be a multiple of the bank size");
    if (nlist_len < 0) throw arg_error("negative neighbour list length");
    // validate the maps and neighbour indices on the host
    for (int v = 0; v < n; ++v) {
        int sl = real_map[v];
        if (sl < 0 || sl >= M || part_map[sl] != v) throw arg_error("MAS real_map/part_map inconsistent");
        if ((long long)nstart[v] + nnum[v] > nlist_len) throw arg_error("MAS neighbour CSR out of range");
    }
    for (int i = 0; i < nlist_len; ++i)
        if ((int)nlist[i] >= n) throw arg_error("MAS neighbour index out of range");
    for (int s2 = 0; s2 < M; ++s2)
        if (part_map[s2] >= n) throw arg_error("MAS part_map entry out of range");
    MasState& m = c.mas;
    m.ready = false;
    m.bank = bank;
    m.n = n;
    m.slot_ordered = true;  // vertex ids increase with their bank slot
    for (int v = 1; v < n; ++v)
        if (real_map[v] <= real_map[v - 1]) { m.slot_ordered = false; break; }
    m.node_offset = node_offset;
    m.M = M;
    m.L = num_levels(std::max(M, n), bank);
    int Mpad = (M + bank - 1) / bank * bank;
    m.going_cap = (size_t)Mpad + (size_t)m.L * ((size_t)n + bank) + bank;
    m.part_map.reserve((size_t)M + 1);
    m.real_map.reserve((size_t)n + 1);
    m.nstart.reserve((size_t)n + 1);
    m.nnum.reserve((size_t)n + 1);
    m.nlist.reserve((size_t)nlist_len + 1);
    if (M) SGIPC_HIP(hipMemcpy(m.part_map.p, part_map, sizeof(int) * M, hipMemcpyHostToDevice));
    if (n) {
        SGIPC_HIP(hipMemcpy(m.real_map.p, real_map, sizeof(int) * n, hipMemcpyHostToDevice));
        SGIPC_HIP(hipMemcpy(m.nstart.p, nstart, sizeof(unsigned) * n, hipMemcpyHostToDevice));
        SGIPC_HIP(hipMemcpy(m.nnum.p, nnum, sizeof(unsigned) * n, hipMemcpyHostToDevice));
    }
    if (nlist_len) SGIPC_HIP(hipMemcpy(m.nlist.p, nlist, sizeof(unsigned) * nlist_len, hipMemcpyHostToDevice));
    m.fine_mask.reserve((size_t)n + 1);
    m.next_mask.reserve((size_t)n + bank + 1);
    size_t nbk = (size_t)std::max(M, n) / bank + 2;
    m.prefix0.reserve(nbk);
    m.psum0.reserve(nbk + 1);
    m.next_prefix.reserve(nbk);
    m.next_psum.reserve(nbk + 1);
    m.level_size.reserve(2 * ((size_t)m.L + 2));
    m.coarse_space.reserve((size_t)m.L * n + 1);
    m.going_next.reserve(m.going_cap);
    m.comp_mask.reserve(m.going_cap);
    m.coarse_table.reserve(6 * (size_t)n + 1);
    m.node_of.reserve((size_t)m.L * n + 1);
    m.level_size_h.assign(2 * ((size_t)m.L + 2), 0);
}

void mas_build(Ctx& c, int sub_offset, int sub_nb, const int* pairs, int npairs) {
    MasState& m = c.mas;
    if (!c.have_bcoo) throw state_error("MAS build: convert the system first");
    if (m.M == 0 && m.n == 0) throw state_error("MAS build: topology not set");
    hipStream_t s = c.stream;
    const int n = m.n, M = m.M, B = m.bank, L = m.L;
    const int Mpad = (M + B - 1) / B * B;
    if (sub_nb != n) throw arg_error("MAS build: subsystem block count differs from the MAS topology");
    m.ready = false;
    int* ls = m.level_size.p;
    fill<int>(ls, 0, 2LL * (L + 2), s);
    fill<int>(m.going_next.p, 0, (long long)m.going_cap, s);
    // level 0
    hipLaunchKernelGGL(k_l0_masks, dim3(div_up(M, 256)), dim3(256), 0, s, M, B, m.part_map.p, m.real_map.p, m.nstart.p,
                       m.nnum.p, m.nlist.p, m.fine_mask.p);
    SGIPC_LAUNCH_CHECK();
    if (npairs > 0) {
        hipLaunchKernelGGL(k_collision, dim3(div_up(npairs, 256)), dim3(256), 0, s, npairs, (const int4*)pairs,
                           m.node_offset, n, B, -1, m.real_map.p, m.coarse_space.p, m.fine_mask.p);
        SGIPC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_l0_closure, dim3(div_up(M, kTB)), dim3(kTB), 0, s, M, B, m.part_map.p, m.fine_mask.p,
                       m.prefix0.p);
    SGIPC_LAUNCH_CHECK();
    int* tot = c.dev_int.reserve(8);
    exclusive_scan(c, m.prefix0.p, m.psum0.p, M / B, tot + 6, s);
    hipLaunchKernelGGL(k_set_level_size, dim3(1), dim3(1), 0, s, ls, 0, tot + 6, B, Mpad);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_level1, dim3(div_up(M, kTB)), dim3(kTB), 0, s, M, Mpad, n, B, m.part_map.p, m.fine_mask.p,
                       m.psum0.p, m.coarse_space.p, m.going_next.p);
    SGIPC_LAUNCH_CHECK();
    const int nmax = n;  // level-local node count never exceeds n
    for (int level = 1; level < L; ++level) {
        fill<unsigned>(m.next_mask.p, 0u, (long long)nmax + B, s);
        hipLaunchKernelGGL(k_lx_masks, dim3(div_up(n, 256)), dim3(256), 0, s, n, B, level, m.nstart.p, m.nnum.p,
                           m.nlist.p, m.coarse_space.p, m.next_mask.p);
        SGIPC_LAUNCH_CHECK();
        if (npairs > 0) {
            hipLaunchKernelGGL(k_collision, dim3(div_up(npairs, 256)), dim3(256), 0, s, npairs, (const int4*)pairs,
                               m.node_offset, n, B, level, m.real_map.p, m.coarse_space.p, m.next_mask.p);
            SGIPC_LAUNCH_CHECK();
        }
        int nbk = div_up(nmax, B);
        hipLaunchKernelGGL(k_lx_closure, dim3(div_up((long long)nbk * B, kTB)), dim3(kTB), 0, s, nbk * B, B, level, ls,
                           m.next_mask.p, m.next_prefix.p, m.comp_mask.p);
        SGIPC_LAUNCH_CHECK();
        exclusive_scan(c, m.next_prefix.p, m.next_psum.p, nbk, tot + 6, s);
        hipLaunchKernelGGL(k_set_level_size, dim3(1), dim3(1), 0, s, ls, level, tot + 6, B, 0);
        SGIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_lx_renumber, dim3(div_up((long long)nbk * B, kTB)), dim3(kTB), 0, s, B, level, ls,
                           m.next_mask.p, m.next_psum.p, m.going_next.p, m.going_cap);
        SGIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_next_table, dim3(div_up(n, 256)), dim3(256), 0, s, n, level, m.coarse_space.p,
                           m.next_mask.p);
        SGIPC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_aggregate, dim3(div_up(n, 256)), dim3(256), 0, s, n, L, m.real_map.p, m.going_next.p,
                       m.coarse_table.p, m.node_of.p);
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemcpyAsync(m.level_size_h.data(), ls, sizeof(int) * 2 * (L + 1), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    m.total = m.level_size_h[2 * L + 1];
    if (m.total <= 0 || m.total % B) throw state_error("MAS build: inconsistent cluster count");
    m.nbanks = m.total / B;
    const int npack = B * (B + 1) / 2;
    const int D = 3 * B;
    m.bank_upper.reserve((size_t)m.nbanks * npack * 9);
    m.inv64.reserve((size_t)m.nbanks * (D * (D + 1) / 2));
    m.inv32.reserve((size_t)m.nbanks * (D * (D + 1) / 2));
    // level-0 bank ranges of the level-1 banks (fused restriction)
    {
        const int nb0 = M / B;
        m.nb1 = div_up(m.level_size_h[2], B);
        int* lo = m.l1lo.reserve((size_t)m.nb1 + 1);
        int* hi = m.l1hi.reserve((size_t)m.nb1 + 1);
        fill<int>(lo, nb0, m.nb1 + 1, s);
        fill<int>(hi, 0, m.nb1 + 1, s);
        if (nb0 > 0) {
            hipLaunchKernelGGL(k_l1_ranges, dim3(div_up(nb0, 256)), dim3(256), 0, s, nb0, B, m.prefix0.p, m.psum0.p,
                               lo, hi);
            SGIPC_LAUNCH_CHECK();
        }
    }
    fill<double>(m.bank_upper.p, 0.0, (long long)m.nbanks * npack * 9, s);
    // The banked projection (16-vertex banks, vertex ids in slot order) sums
    // coarse levels in LDS with few atomics; every other case -- and the
    // deterministic mode -- uses the ordered projection, which sorts
    // (entry, block) records and sums each entry in a fixed order: no atomic
    // contention where many fine blocks land on few coarse entries
    // (32-vertex banks: 23 ms with global FP64 atomics at 137 k vertices).
    const bool banked = (B == 16) && m.slot_ordered && m.nb1 > 0 && L >= 2;
    if (c.K > 0 && (c.deterministic || !banked)) {
        if ((long long)c.K >= (1LL << 29)) throw capacity_error("MAS build: too many blocks for the ordered projection");
        const unsigned sentinel = (unsigned)((long long)m.nbanks * npack);
        int bits = 1;
        while (bits < 32 && (1ULL << bits) <= (unsigned long long)sentinel) ++bits;
        const long long mrec = (long long)c.K * std::max(L - 1, 0);
        if (L > 1) {
            unsigned* key = m.proj_key.reserve(2 * (size_t)mrec);
            unsigned* rec = m.proj_rec.reserve(2 * (size_t)mrec);
            hipLaunchKernelGGL(k_proj_records, dim3(div_up(c.K, 256)), dim3(256), 0, s, c.K, (const int*)c.t_rows,
                               (const int*)c.t_cols, (const double*)c.t_vals, sub_offset, n, B, L,
                               (const int*)m.node_of.p, m.bank_upper.p, sentinel, key, rec);
            SGIPC_LAUNCH_CHECK();
            if (c.sort_impl == SORT_OWN) {
                radix_sort_pairs(c, key, key + mrec, (const int*)rec, 0, (int*)(rec + mrec), mrec, bits, s);
            } else {  // rocPRIM as the cross-check
                size_t bytes = 0;
                SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key, key + mrec, rec, rec + mrec, (size_t)mrec, 0,
                                                    bits, s));
                void* tmp = m.proj_sort_tmp.reserve(bytes + 1);
                SGIPC_HIP(rocprim::radix_sort_pairs(tmp, bytes, key, key + mrec, rec, rec + mrec, (size_t)mrec, 0,
                                                    bits, s));
            }
            double* contrib = m.proj_contrib.reserve(9 * (size_t)mrec);
            const unsigned* skey = key + mrec;
            hipLaunchKernelGGL(k_proj_contrib, dim3(div_up(mrec, 256)), dim3(256), 0, s, mrec, skey,
                               (const unsigned*)(rec + mrec), (const int*)c.t_rows, (const int*)c.t_cols,
                               (const double*)c.t_vals, sub_offset, n, B, (const int*)m.node_of.p, sentinel, contrib);
            SGIPC_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_proj_pieces, dim3(div_up(mrec, 256)), dim3(256), 0, s, mrec, skey, sentinel, contrib);
            SGIPC_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_proj_entries, dim3(div_up(mrec, 256)), dim3(256), 0, s, mrec, skey, sentinel,
                               (const double*)contrib, m.bank_upper.p);
            SGIPC_LAUNCH_CHECK();
        } else {
            hipLaunchKernelGGL(k_proj_records, dim3(div_up(c.K, 256)), dim3(256), 0, s, c.K, (const int*)c.t_rows,
                               (const int*)c.t_cols, (const double*)c.t_vals, sub_offset, n, B, L,
                               (const int*)m.node_of.p, m.bank_upper.p, sentinel, (unsigned*)nullptr,
                               (unsigned*)nullptr);
            SGIPC_LAUNCH_CHECK();
        }
    } else if (c.K > 0 && banked) {
        // top levels: the highest levels whose banks fit kProjTopBanksMax
        // (the kernel with kProjTopBanks of them, when they fit, runs two
        // workgroups per CU; a level left out goes to global atomics)
        int top_lvl = L, top_lo = m.nbanks, top_nb = 0;
        for (int l = L - 1; l >= 1; --l) {
            int beg = m.level_size_h[2 * l + 1] / B;
            int end = m.level_size_h[2 * (l + 1) + 1] / B;
            if (top_nb + (end - beg) > kProjTopBanksMax) break;
            top_nb += end - beg;
            top_lo = beg;
            top_lvl = l;
        }
        const int win_levels = std::min(kProjWinLevels, top_lvl - 1);
        const int nb0 = M / B;
        int* vlo = m.own_lo.reserve((size_t)m.nb1 + 1);
        int* vhi = m.own_hi.reserve((size_t)m.nb1 + 1);
        fill<int>(vlo, 0x7fffffff, m.nb1, s);
        fill<int>(vhi, 0, m.nb1, s);
        hipLaunchKernelGGL(k_owned_vertices, dim3(div_up(nb0, 256)), dim3(256), 0, s, nb0, B, m.prefix0.p, m.psum0.p,
                           m.part_map.p, vlo, vhi);
        SGIPC_LAUNCH_CHECK();
        const int top_len = top_nb * kProjBank16;
        double* scratch = m.proj_scratch.reserve((size_t)m.nb1 * (top_len > 0 ? top_len : 1));
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(m.nb1), dim3(512), 0, s, (const int*)c.uniq_start.p, (const int*)c.t_rows,
                               (const int*)c.t_cols, (const double*)c.t_vals, sub_offset, n, L,
                               (const int*)m.node_of.p, (const int*)vlo, (const int*)vhi, m.nbanks, m.bank_upper.p,
                               win_levels, top_lvl, top_lo, top_nb, scratch);
        };
        if (top_nb <= kProjTopBanks) launch(k_project_banked<kProjTopBanks>);
        else launch(k_project_banked<kProjTopBanksMax>);
        SGIPC_LAUNCH_CHECK();
        if (top_len > 0) {
            hipLaunchKernelGGL(k_project_top_reduce, dim3(div_up(top_len, 64)), dim3(256), 0, s,
                               (const double*)scratch, m.nb1, top_len, m.bank_upper.p + (size_t)top_lo * kProjBank16);
            SGIPC_LAUNCH_CHECK();
        }
    }
    // the FP32-with-refinement inverses serve FP32 applies only
    int inv_mode = c.bank_inverse;
    if (inv_mode == BANK_INV_MFMA32_REFINED && c.precond_kind != 3) inv_mode = BANK_INV_MFMA64;
    m.inv64_ok = inv_mode != BANK_INV_MFMA32_REFINED;
    // half storage for the fused FP32 bank solves (16-vertex banks only):
    // written by the FP64 MFMA inversion itself, else packed afterwards
    m.inv16_ok = false;
    const bool half = c.bank_store == BANK_STORE_F16 && B == 16 && m.nbanks > 0;
    if (half) {
        m.inv16.reserve((size_t)m.nbanks * (D * (D + 1) / 2) + 8);  // 2352 bytes a bank: 16-byte aligned banks
        m.inv16_d.reserve((size_t)m.nbanks * D);
    }
    const bool half_fused = half && inv_mode == BANK_INV_MFMA64;
    if (inv_mode != BANK_INV_VALU64) {
        bank_inverse_mfma(c, B, m.nbanks, m.bank_upper.p, m.inv64.p, m.inv32.p, inv_mode,
                          half_fused ? m.inv16.p : nullptr, half_fused ? m.inv16_d.p : nullptr);
    } else {
        if (B == 16)
            hipLaunchKernelGGL(k_bank_inverse_reg48, dim3(m.nbanks), dim3(64), 0, s, m.nbanks, m.bank_upper.p,
                               m.inv64.p, m.inv32.p);
        else
            hipLaunchKernelGGL(k_bank_inverse<32>, dim3(m.nbanks), dim3(128), 0, s, m.nbanks, m.bank_upper.p,
                               m.inv64.p, m.inv32.p);
        SGIPC_LAUNCH_CHECK();
    }
    if (half && !half_fused) {
        if (m.inv64_ok)
            hipLaunchKernelGGL(k_pack_inv_half<double>, dim3(m.nbanks), dim3(64), 0, s, m.nbanks,
                               (const double*)m.inv64.p, m.inv16.p, m.inv16_d.p);
        else
            hipLaunchKernelGGL(k_pack_inv_half<float>, dim3(m.nbanks), dim3(64), 0, s, m.nbanks,
                               (const float*)m.inv32.p, m.inv16.p, m.inv16_d.p);
        SGIPC_LAUNCH_CHECK();
    }
    m.inv16_ok = half;
    // apply work vectors: allocated here so their addresses are stable for
    // captured PCG graphs; padding entries stay zero (never written)
    m.R64.reserve(3 * (size_t)m.total + 3);
    m.Z64.reserve(3 * (size_t)m.total + 3);
    m.R32.reserve(3 * (size_t)m.total + 3);
    m.Z32.reserve(3 * (size_t)m.total + 3);
    fill<double>(m.R64.p, 0.0, 3LL * m.total + 3, s);
    fill<float>(m.R32.p, 0.0f, 3LL * m.total + 3, s);
    m.ready = true;
}

void mas_apply64(Ctx& c, const double* r, double* z) {
    if (!c.mas.ready) throw state_error("MAS apply: build first");
    if (!c.mas.inv64_ok) throw state_error("MAS apply: only FP32 inverses were built");
    mas_apply_t<double>(c, r, z, c.mas.R64.p, c.mas.Z64.p, c.mas.inv64.p);
}
void mas_apply32(Ctx& c, const double* r, double* z) {
    if (!c.mas.ready) throw state_error("MAS apply: build first");
    mas_apply_t<float>(c, r, z, c.mas.R32.p, c.mas.Z32.p, c.mas.inv32.p);
}
void mas_restrict(Ctx& c, const double* r, long long sub3, bool fp32, const int* ctrl, bool top, SceneSkip skip) {
    if (fp32) mas_restrict_t<double, float>(c, r, sub3, c.mas.R32.p, ctrl, top, skip);
    else mas_restrict_t<double, double>(c, r, sub3, c.mas.R64.p, ctrl, top, skip);
}
void mas_restrict(Ctx& c, const float* r, long long sub3, bool fp32, const int* ctrl, bool top, SceneSkip skip) {
    if (!fp32) throw state_error("MAS: FP32 residual needs the FP32 preconditioner data");
    mas_restrict_t<float, float>(c, r, sub3, c.mas.R32.p, ctrl, top, skip);
}


// workgroups of the fused update: level-1 banks (grid-stride) plus a few for
// the dofs outside the preconditioned range; one r.r partial each
void mas_update_layout(const Ctx& c, int& g1, int& gx) {
    const MasState& m = c.mas;
    long long nout = 3LL * c.nb - 3LL * m.n;
    gx = nout > 0 ? (int)std::min<long long>(64, div_up(nout, kTB)) : 0;
    g1 = std::max(1, std::min(m.nb1, kPart - gx));
}

template <typename V>
static void update_restrict_t(Ctx& c, double* x, V* r, const V* p, const V* Ap, const double* part_pap, int np_pap,
                              const double* part_rz, int np_rz, double* part_rr, int* ctrl, bool fp32, bool top) {
    MasState& m = c.mas;
    int g1, gx;
    mas_update_layout(c, g1, gx);
    const long long N = 3LL * c.nb;
    if (fp32) {
        hipLaunchKernelGGL((k_update_restrict<V, float>), dim3(g1 + gx), dim3(kTB), 0, c.stream, N, x, r, p, Ap,
                           part_pap, np_pap, part_rz, np_rz, part_rr, ctrl, c.sub_offset, m.n, m.bank, m.L,
                           (const int*)m.level_size.p, (const int*)m.part_map.p, (const int*)m.going_next.p,
                           (const unsigned*)m.fine_mask.p, (const unsigned*)m.comp_mask.p,
                           (const int*)m.l1lo.p, (const int*)m.l1hi.p, m.nb1, g1, m.R32.p);
    } else {
        if (sizeof(V) != 8) throw state_error("MAS: FP32 residual needs the FP32 preconditioner data");
        hipLaunchKernelGGL((k_update_restrict<V, double>), dim3(g1 + gx), dim3(kTB), 0, c.stream, N, x, r, p, Ap,
                           part_pap, np_pap, part_rz, np_rz, part_rr, ctrl, c.sub_offset, m.n, m.bank, m.L,
                           (const int*)m.level_size.p, (const int*)m.part_map.p, (const int*)m.going_next.p,
                           (const unsigned*)m.fine_mask.p, (const unsigned*)m.comp_mask.p,
                           (const int*)m.l1lo.p, (const int*)m.l1hi.p, m.nb1, g1, m.R64.p);
    }
    SGIPC_LAUNCH_CHECK();
    if (m.L >= 4 && top) {
        if (fp32)
            hipLaunchKernelGGL((k_restrict_top<float>), dim3(1), dim3(kTop), 0, c.stream, m.bank, m.L,
                               (const int*)m.level_size.p, (const int*)m.going_next.p, (const unsigned*)m.comp_mask.p,
                               m.R32.p, (const int*)ctrl);
        else
            hipLaunchKernelGGL((k_restrict_top<double>), dim3(1), dim3(kTop), 0, c.stream, m.bank, m.L,
                               (const int*)m.level_size.p, (const int*)m.going_next.p, (const unsigned*)m.comp_mask.p,
                               m.R64.p, (const int*)ctrl);
        SGIPC_LAUNCH_CHECK();
    }
}

void mas_update_restrict(Ctx& c, double* x, double* r, const double* p, const double* Ap, const double* part_pap,
                         int np_pap, const double* part_rz, int np_rz, double* part_rr, int* ctrl, bool fp32,
                         bool top) {
    update_restrict_t<double>(c, x, r, p, Ap, part_pap, np_pap, part_rz, np_rz, part_rr, ctrl, fp32, top);
}
void mas_update_restrict(Ctx& c, double* x, float* r, const float* p, const float* Ap, const double* part_pap,
                         int np_pap, const double* part_rz, int np_rz, double* part_rr, int* ctrl, bool fp32,
                         bool top) {
    update_restrict_t<float>(c, x, r, p, Ap, part_pap, np_pap, part_rz, np_rz, part_rr, ctrl, fp32, top);
}

template <typename T, bool H>
static void solve_top_t(Ctx& c, const void* inv, const float* dsc, T* R, T* Z, const int* ctrl, SceneSkip skip) {
    MasState& m = c.mas;
    const int top_bank0 = m.L >= 4 ? m.level_size_h[7] / m.bank : m.nbanks;
    prof_begin(c, PROF_MAS_SOLVE);
    hipLaunchKernelGGL((k_mas_solve_top<T, H>), dim3(1 + div_up(top_bank0, kCW)), dim3(kCT), 0, c.stream, m.L,
                       (const int*)m.level_size.p, (const int*)m.going_next.p, (const unsigned*)m.comp_mask.p,
                       top_bank0, m.nbanks, inv, dsc, R, Z, ctrl, skip);
    SGIPC_LAUNCH_CHECK();
    prof_end(c, PROF_MAS_SOLVE);
}

bool mas_fused(const Ctx& c) { return c.mas_fused && c.mas.bank == 16 && c.mas.ready; }

void mas_solve_with_top(Ctx& c, bool fp32, const int* ctrl, SceneSkip skip) {
    MasState& m = c.mas;
    if (fp32 && m.inv16_ok) solve_top_t<float, true>(c, m.inv16.p, m.inv16_d.p, m.R32.p, m.Z32.p, ctrl, skip);
    else if (fp32) solve_top_t<float, false>(c, m.inv32.p, nullptr, m.R32.p, m.Z32.p, ctrl, skip);
    else {
        if (!m.inv64_ok) throw state_error("MAS: only FP32 inverses were built");
        solve_top_t<double, false>(c, m.inv64.p, nullptr, m.R64.p, m.Z64.p, ctrl, skip);
    }
}

void mas_bank_solve(Ctx& c, bool fp32, const int* ctrl, SceneSkip skip) {
    if (fp32) mas_solve_t<float>(c, c.mas.inv32.p, c.mas.R32.p, c.mas.Z32.p, ctrl, skip);
    else {
        if (!c.mas.inv64_ok) throw state_error("MAS: only FP32 inverses were built");
        mas_solve_t<double>(c, c.mas.inv64.p, c.mas.R64.p, c.mas.Z64.p, ctrl, skip);
    }
}

}  // namespace sgipc
