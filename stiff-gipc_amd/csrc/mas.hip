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
__global__ void k_l0_masks(int M, int B, const int* __restrict__ part_map, const int* __restrict__ real_map,
                           const unsigned* __restrict__ nstart, const unsigned* __restrict__ nnum,
                           const unsigned* __restrict__ nlist, unsigned* __restrict__ fine_mask) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    int v = part_map[s];
    if (v < 0) return;
    int bank = s / B, lane = s % B;
    unsigned msk = 1u << lane, st = nstart[v], nn = nnum[v];
    for (unsigned i = 0; i < nn; ++i) {
        int su = real_map[nlist[st + i]];
        if (su / B == bank) msk |= 1u << (su % B);
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
    if (p >= np) return;
    int v[4], cp[4];
    int cnt = decode_pair(pairs[p], node_offset, v);
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
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (cp[i] < 0 || msk[i] == 0u) continue;
        atomicOr(dst + (level < 0 ? v[i] : cp[i]), msk[i]);
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
        for (unsigned i = 0; i < nn; ++i) {
            int cu = cs[nlist[st + i]];
            if (cu / B == ci / B) msk |= 1u << (cu % B);
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
                                                     unsigned* __restrict__ next_mask, int* __restrict__ next_prefix) {
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
    if (valid) next_mask[c] = cl;
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

// Galerkin projection of every BCOO block into the bank matrices of all
// levels.  Level 0 targets are unique (plain stores); coarse levels whose
// banks all fit in LDS (the top levels, where many fine blocks land on few
// entries) are pre-summed per workgroup with LDS atomics and flushed once;
// the remaining coarse levels use global FP64 atomics.
constexpr int kProjLdsDoubles = 6 * 1224;  // 6 banks of 16 (136 blocks x 9) or 1 bank of 32 + margin
__global__ void __launch_bounds__(256) k_project(int K, const int* __restrict__ rows, const int* __restrict__ cols,
                                                 const double* __restrict__ vals, int off, int n, int B, int L,
                                                 const int* __restrict__ node_of, double* __restrict__ bank_upper,
                                                 int lds_bank_lo, int lds_nbanks) {
    __shared__ double acc[kProjLdsDoubles];
    const int npack = B * (B + 1) / 2;
    const int lds_len = lds_nbanks * npack * 9;
    for (int i = threadIdx.x; i < lds_len; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x) {
        int r = rows[k] - off, c = cols[k] - off;
        if (r < 0 || c < 0 || r >= n || c >= n) continue;
        double H[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) H[e] = vals[9 * (size_t)k + e];  // column-major
        for (int l = 0; l < L; ++l) {
            int I = node_of[(size_t)l * n + r], J = node_of[(size_t)l * n + c];
            if (I / B != J / B) continue;
            int li = I % B, lj = J % B, bank = I / B;
            int pidx;
            double blk[9];
            if (li == lj) {
                pidx = pack_index(B, li, li);
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        blk[p + 3 * q] = (r == c) ? H[p + 3 * q] : H[p + 3 * q] + H[q + 3 * p];
            } else if (li < lj) {
                pidx = pack_index(B, li, lj);
#pragma unroll
                for (int e = 0; e < 9; ++e) blk[e] = H[e];
            } else {
                pidx = pack_index(B, lj, li);
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int q = 0; q < 3; ++q) blk[p + 3 * q] = H[q + 3 * p];
            }
            if (l == 0) {
                double* dst = bank_upper + ((size_t)bank * npack + pidx) * 9;  // unique per BCOO block
#pragma unroll
                for (int e = 0; e < 9; ++e) dst[e] += blk[e];
            } else if (bank >= lds_bank_lo && bank < lds_bank_lo + lds_nbanks) {
                double* dst = acc + ((size_t)(bank - lds_bank_lo) * npack + pidx) * 9;
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    if (blk[e] != 0.0) atomicAdd(dst + e, blk[e]);
            } else {
                double* dst = bank_upper + ((size_t)bank * npack + pidx) * 9;
#pragma unroll
                for (int e = 0; e < 9; ++e)
                    if (blk[e] != 0.0) atomicAdd(dst + e, blk[e]);
            }
        }
    }
    __syncthreads();
    double* g = bank_upper + (size_t)lds_bank_lo * npack * 9;
    for (int i = threadIdx.x; i < lds_len; i += blockDim.x)
        if (acc[i] != 0.0) atomicAdd(g + i, acc[i]);
}

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
    // symmetrise: upper triangle mirrored; store dense row-major
    if (i < D) {
        double* o64 = inv64 + (size_t)b * D * D;
        float* o32 = inv32 + (size_t)b * D * D;
        for (int r = 0; r < D; ++r) {
            double v = (r <= i) ? S[r][i] : S[i][r];
            o64[(size_t)r * D + i] = v;
            o32[(size_t)r * D + i] = (float)v;
        }
    }
}

// ---- apply -----------------------------------------------------------------
template <typename T>
__global__ void k_restrict0(int M, const int* __restrict__ part_map, const double* __restrict__ r, T* __restrict__ R) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    int v = part_map[s];
#pragma unroll
    for (int k = 0; k < 3; ++k) R[3 * (size_t)s + k] = (v >= 0) ? (T)r[3 * (size_t)v + k] : (T)0;
}

// Aggregate level-(l-1) nodes into their parents (component sums inside each
// bank, lowest lane of the component writes; no atomics).
// nodes: global ids [beg, beg+cnt) (level 0: slots [0, M) with parent via
// going_next[vertex]).
template <typename T>
__global__ void __launch_bounds__(kTB) k_restrict_up(int level_from, int B, const int* __restrict__ ls, int M,
                                                      const int* __restrict__ part_map,
                                                      const int* __restrict__ going_next, T* __restrict__ R,
                                                      const int* __restrict__ ctrl) {
    if (ctrl && ctrl[0]) return;
    __shared__ int par[kTB];
    __shared__ T sv[3][kTB];
    int c = blockIdx.x * blockDim.x + threadIdx.x;  // local index within level
    int lane = threadIdx.x % B, bbase = threadIdx.x - lane;
    int cnt, beg;
    if (level_from == 0) { cnt = M; beg = 0; }
    else { cnt = ls[2 * level_from + 0]; beg = ls[2 * level_from + 1]; }
    int p = -1;
    T v0 = 0, v1 = 0, v2 = 0;
    if (c < cnt) {
        size_t g = (size_t)beg + c;
        if (level_from == 0) {
            int vert = part_map[c];
            p = (vert >= 0) ? going_next[vert] : -1;
        } else {
            p = going_next[g];
        }
        v0 = R[3 * g]; v1 = R[3 * g + 1]; v2 = R[3 * g + 2];
    }
    par[threadIdx.x] = p;
    sv[0][threadIdx.x] = v0; sv[1][threadIdx.x] = v1; sv[2][threadIdx.x] = v2;
    __syncthreads();
    if (p < 0) return;
    for (int l = 0; l < lane; ++l)
        if (par[bbase + l] == p) return;  // not the lowest lane of its component
    T s0 = 0, s1 = 0, s2 = 0;
    for (int l = lane; l < B; ++l)
        if (par[bbase + l] == p) { s0 += sv[0][bbase + l]; s1 += sv[1][bbase + l]; s2 += sv[2][bbase + l]; }
    R[3 * (size_t)p] = s0; R[3 * (size_t)p + 1] = s1; R[3 * (size_t)p + 2] = s2;
}

template <int B, typename T>
__global__ void __launch_bounds__(B == 16 ? 64 : 128) k_bank_solve(const T* __restrict__ inv, const T* __restrict__ R,
                                                                    T* __restrict__ Z, const int* __restrict__ ctrl) {
    if (ctrl && ctrl[0]) return;
    constexpr int D = 3 * B;
    __shared__ T rs[D];
    int b = blockIdx.x, i = threadIdx.x;
    if (i < D) rs[i] = R[(size_t)b * D + i];
    __syncthreads();
    if (i >= D) return;
    const T* A = inv + (size_t)b * D * D;
    T acc = 0;
#pragma unroll 8
    for (int j = 0; j < D; ++j) acc += A[(size_t)j * D + i] * rs[j];  // symmetric: column i == row i
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

// restriction to the coarse levels + bank solves on all levels; R level 0
// must already hold the slot-ordered residual.  `ctrl` (device, may be
// null) skips the work once the PCG loop has stopped.
template <typename T>
void mas_up_solve_t(Ctx& c, T* R, T* Z, const T* inv, const int* ctrl) {
    MasState& m = c.mas;
    hipStream_t s = c.stream;
    for (int l = 0; l + 1 < m.L; ++l) {
        int cnt = (l == 0) ? m.M : m.level_size_h[2 * l];
        if (cnt <= 0) break;
        hipLaunchKernelGGL(k_restrict_up<T>, dim3(div_up(cnt, kTB)), dim3(kTB), 0, s, l, m.bank, m.level_size.p, m.M,
                           m.part_map.p, m.going_next.p, R, ctrl);
        SGIPC_LAUNCH_CHECK();
    }
    prof_begin(c, PROF_MAS_SOLVE);
    if (m.bank == 16)
        hipLaunchKernelGGL((k_bank_solve<16, T>), dim3(m.nbanks), dim3(64), 0, s, inv, R, Z, ctrl);
    else
        hipLaunchKernelGGL((k_bank_solve<32, T>), dim3(m.nbanks), dim3(128), 0, s, inv, R, Z, ctrl);
    SGIPC_LAUNCH_CHECK();
    prof_end(c, PROF_MAS_SOLVE);
}

template <typename T>
void mas_apply_t(Ctx& c, const double* r, double* z, T* R, T* Z, const T* inv) {
    MasState& m = c.mas;
    hipStream_t s = c.stream;
    if (m.n <= 0) return;
    hipLaunchKernelGGL(k_restrict0<T>, dim3(div_up(m.M, 256)), dim3(256), 0, s, m.M, m.part_map.p, r, R);
    SGIPC_LAUNCH_CHECK();
    mas_up_solve_t<T>(c, R, Z, inv, nullptr);
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
                           m.next_mask.p, m.next_prefix.p);
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
    m.inv64.reserve((size_t)m.nbanks * D * D);
    m.inv32.reserve((size_t)m.nbanks * D * D);
    fill<double>(m.bank_upper.p, 0.0, (long long)m.nbanks * npack * 9, s);
    if (c.K > 0) {
        // coarse levels whose banks fit the LDS accumulator (top levels first)
        int lds_lo = m.nbanks, lds_nb = 0;
        for (int l = L - 1; l >= 1; --l) {
            int beg = m.level_size_h[2 * l + 1] / B;
            int end = (l + 1 < L + 1) ? m.level_size_h[2 * (l + 1) + 1] / B : m.nbanks;
            int nb_l = end - beg;
            if ((lds_nb + nb_l) * npack * 9 > kProjLdsDoubles) break;
            lds_nb += nb_l;
            lds_lo = beg;
        }
        int grid = std::min(div_up(c.K, 256), 1024);
        hipLaunchKernelGGL(k_project, dim3(grid), dim3(256), 0, s, c.K, c.t_rows, c.t_cols, c.t_vals, sub_offset, n, B,
                           L, m.node_of.p, m.bank_upper.p, lds_lo, lds_nb);
        SGIPC_LAUNCH_CHECK();
    }
    if (B == 16)
        hipLaunchKernelGGL(k_bank_inverse<16>, dim3(m.nbanks), dim3(64), 0, s, m.nbanks, m.bank_upper.p, m.inv64.p,
                           m.inv32.p);
    else
        hipLaunchKernelGGL(k_bank_inverse<32>, dim3(m.nbanks), dim3(128), 0, s, m.nbanks, m.bank_upper.p, m.inv64.p,
                           m.inv32.p);
    SGIPC_LAUNCH_CHECK();
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
    mas_apply_t<double>(c, r, z, c.mas.R64.p, c.mas.Z64.p, c.mas.inv64.p);
}
void mas_apply32(Ctx& c, const double* r, double* z) {
    if (!c.mas.ready) throw state_error("MAS apply: build first");
    mas_apply_t<float>(c, r, z, c.mas.R32.p, c.mas.Z32.p, c.mas.inv32.p);
}
void mas_up_solve64(Ctx& c, const int* ctrl) { mas_up_solve_t<double>(c, c.mas.R64.p, c.mas.Z64.p, c.mas.inv64.p, ctrl); }
void mas_up_solve32(Ctx& c, const int* ctrl) { mas_up_solve_t<float>(c, c.mas.R32.p, c.mas.Z32.p, c.mas.inv32.p, ctrl); }

}  // namespace sgipc
