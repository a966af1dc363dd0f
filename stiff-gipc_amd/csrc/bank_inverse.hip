// Bank inverses of the multilevel additive Schwarz preconditioner on the
// matrix cores.
//
// Reference: __inverse6_P96x96 (S/MASPreconditioner.cu:630-726), called from
// MASPreconditioner::PrepareHessian_bcoo (:1816-2093): unpivoted scalar
// Gauss-Jordan of every 3B x 3B bank matrix in FP64 shared memory (a zero
// diagonal entry replaced by 1), the inverse stored in FP32.
//
// Here the same unpivoted elimination runs over 16 x 16 tiles, one
// wavefront per bank, the matrix in LDS.  Step k inverts the diagonal tile
// P = A_kk in registers (scalar Gauss-Jordan across the wavefront) and then
//   A_kj <- P^-1 A_kj,  A_ij <- A_ij - A_ik A_kj,  A_ik <- -A_ik P^-1,
//   A_kk <- P^-1                                  (i, j != k)
// with every tile product on MFMA (v_mfma_f64_16x16x4_f64, or
// v_mfma_f32_16x16x4_f32 for the FP32 factorisation).  In exact arithmetic
// this is the scalar elimination (both give A^-1); the rounding differs.
//
// Two forms:
//  * FP64: the elimination in FP64 MFMA; writes the FP64 inverse and its FP32
//    rounding (the reference's stored precision).
//  * FP32 + FP64 residual: the elimination in FP32 MFMA gives X0; one
//    refinement step X1 = X0 + X0 (I - A X0) with the residual I - A X0
//    accumulated in FP64 MFMA from the FP64 bank matrix; X1 is stored FP32.
#include "context.hpp"
#include <utility>

namespace sgipc {
namespace {

typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));

// MFMA 16x16x4 lane layouts (l = lane, q = l / 16): operand A chunk c of a
// tile is element (l % 16, 4c + q); operand B chunk c is element
// (4c + q, l % 16); accumulator register r holds column l % 16 and row
// row(q, r): 4q + r for v_mfma_f32_16x16x4_f32, q + 4r for
// v_mfma_f64_16x16x4_f64 (measured on gfx950).  grp(i) / reg(i): the q and
// r that hold row i.
template <typename T>
struct Mf;
template <>
struct Mf<double> {
    typedef dv4 acc;
    __device__ static acc mac(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __host__ __device__ static constexpr int row(int q, int r) { return q + 4 * r; }
    __host__ __device__ static constexpr int grp(int i) { return i & 3; }
    __host__ __device__ static constexpr int reg(int i) { return i >> 2; }
};
template <>
struct Mf<float> {
    typedef fv4 acc;
    __device__ static acc mac(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __host__ __device__ static constexpr int row(int q, int r) { return 4 * q + r; }
    __host__ __device__ static constexpr int grp(int i) { return i >> 2; }
    __host__ __device__ static constexpr int reg(int i) { return i & 3; }
};

// A square matrix in LDS, row-major with stride LD, seen as 16 x 16 tiles,
// read and written in the MFMA operand and accumulator layouts above.
template <typename T, int LD>
struct Tiles {
    T* s;
    typedef typename Mf<T>::acc acc;
    __device__ T& at(int i, int j) const { return s[i * LD + j]; }
    __device__ acc load_c(int ti, int tj, int l) const {
        acc a;
        const int q = l >> 4, j = 16 * tj + (l & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = at(16 * ti + Mf<T>::row(q, r), j);
        return a;
    }
    __device__ void store_c(int ti, int tj, int l, acc a) const {
        const int q = l >> 4, j = 16 * tj + (l & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) at(16 * ti + Mf<T>::row(q, r), j) = a[r];
    }
    __device__ T opa(int ti, int tk, int c, int l) const { return at(16 * ti + (l & 15), 16 * tk + 4 * c + (l >> 4)); }
    __device__ T opb(int tk, int tj, int c, int l) const { return at(16 * tk + 4 * c + (l >> 4), 16 * tj + (l & 15)); }
};

// Inverse of the 16 x 16 tile held in accumulator layout: 16 unpivoted
// Gauss-Jordan steps; the pivot row and column reach every lane by lane
// shuffles.  Step p: a_pp <- 1/a_pp, a_pj <- a_pj / a_pp,
// a_ip <- -a_ip / a_pp, a_ij <- a_ij - a_ip a_pj / a_pp.
template <typename T>
__device__ inline typename Mf<T>::acc invert_tile(typename Mf<T>::acc P, int l) {
    const int q = l >> 4, j = l & 15;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int pg = Mf<T>::grp(p), pr = Mf<T>::reg(p);
        const T rowv = __shfl(P[pr], 16 * pg + j, 64);  // a_pj
        T colv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) colv[r] = __shfl(P[r], 16 * q + p, 64);  // a_ip, i = row(q, r)
        const T piv = __shfl(P[pr], 16 * pg + p, 64);
        const T inv = (T)1 / piv;
        const T rs = rowv * inv;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = Mf<T>::row(q, r);
            T v;
            if (i == p) v = (j == p) ? inv : rs;
            else v = (j == p) ? -colv[r] * inv : P[r] - colv[r] * rs;
            P[r] = v;
        }
    }
    return P;
}

__device__ inline int pack_block(int B, int r, int c) { return B * r - r * (r + 1) / 2 + c; }  // r <= c

// element (i, j) of bank b's matrix from its packed upper 3x3 blocks
// (column-major), a zero diagonal entry read as 1
template <int D>
__device__ inline double bank_elem(const double* __restrict__ src, int i, int j) {
    constexpr int B = D / 3;
    const int ri = i / 3, pi = i % 3, cj = j / 3, qj = j % 3;
    double v = (ri <= cj) ? src[(size_t)pack_block(B, ri, cj) * 9 + pi + 3 * qj]
                          : src[(size_t)pack_block(B, cj, ri) * 9 + qj + 3 * pi];
    if (i == j && v == 0.0) v = 1.0;
    return v;
}

// row r of packed element e of a D x D upper triangle packed by rows
// (row r starts at r (2D - r + 1) / 2)
template <int D>
__device__ inline int packed_row(int e) {
    const float b = 2.0f * D + 1.0f;
    int r = (int)((b - sqrtf(b * b - 8.0f * (float)e)) * 0.5f);
    r = r < 0 ? 0 : (r > D - 1 ? D - 1 : r);
    while (r > 0 && r * (2 * D - r + 1) / 2 > e) --r;
    while (r + 1 < D && (r + 1) * (2 * D - r) / 2 <= e) ++r;
    return r;
}

// In-place blocked Gauss-Jordan of the NT x NT-tile matrix in S (one
// wavefront).  W: a 16 x 16 tile of scratch.
template <typename T, int NT, int LD, int LDW>
__device__ inline void block_gauss_jordan(const Tiles<T, LD>& S, const Tiles<T, LDW>& W, int l) {
    typedef typename Mf<T>::acc acc;
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        acc P = invert_tile<T>(S.load_c(k, k, l), l);
        W.store_c(0, 0, l, P);
        __syncthreads();
        acc Trow[NT], Tcol[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (t == k) continue;
            acc a = {0, 0, 0, 0}, c = {0, 0, 0, 0};
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
                a = Mf<T>::mac(W.opa(0, 0, ch, l), S.opb(k, t, ch, l), a);  // P^-1 A_kt
                c = Mf<T>::mac(S.opa(t, k, ch, l), W.opb(0, 0, ch, l), c);  // A_tk P^-1 (old A_tk)
            }
            Trow[t] = a;
            Tcol[t] = c;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (t != k) S.store_c(k, t, l, Trow[t]);
        __syncthreads();
        // trailing update with the new row tiles and the old column tiles
#pragma unroll
        for (int i = 0; i < NT; ++i) {
            if (i == k) continue;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (t == k) continue;
                acc a = S.load_c(i, t, l);
#pragma unroll
                for (int ch = 0; ch < 4; ++ch) a = Mf<T>::mac(-S.opa(i, k, ch, l), S.opb(k, t, ch, l), a);
                S.store_c(i, t, l, a);
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (t != k) S.store_c(t, k, l, -Tcol[t]);
        S.store_c(k, k, l, P);
        __syncthreads();
    }
}

// FP64 (T = double) or FP32 (T = float) elimination of bank blockIdx.x.
// FP64 writes inv64 and its FP32 rounding inv32; FP32 writes inv32 (X0).
template <typename T, int NT>
__global__ void __launch_bounds__(64) k_bank_inverse_mfma(int nbanks, const double* __restrict__ bank_upper,
                                                          double* __restrict__ inv64, float* __restrict__ inv32) {
    constexpr int D = 16 * NT, B = D / 3, npack = B * (B + 1) / 2, NP = D * (D + 1) / 2;
    constexpr int LD = D + 1, LDW = 17;
    __shared__ T sm[D * LD];
    __shared__ T sw[16 * LDW];
    const Tiles<T, LD> S{sm};
    const Tiles<T, LDW> W{sw};
    const int b = blockIdx.x, l = threadIdx.x;
    const double* src = bank_upper + (size_t)b * npack * 9;
    for (int e = l; e < D * D; e += 64) {
        const int i = e / D, j = e % D;
        S.at(i, j) = (T)bank_elem<D>(src, i, j);
    }
    __syncthreads();
    block_gauss_jordan<T, NT>(S, W, l);
    // upper triangle (r <= c) packed by rows
    for (int e = l; e < NP; e += 64) {
        const int r = packed_row<D>(e);
        const int c = r + (e - r * (2 * D - r + 1) / 2);
        const T v = S.at(r, c);
        if constexpr (sizeof(T) == 8) inv64[(size_t)b * NP + e] = v;
        inv32[(size_t)b * NP + e] = (float)v;
    }
}

// Value of v in lane LANE (a compile-time constant) of the wavefront,
// through scalar registers.
template <typename T>
__device__ __forceinline__ T read_lane(T v, int lane) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffu), lane);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
    }
}

// Step J of the scalar Gauss-Jordan of a 16 x 16 tile held column per lane
// (lane i < 16 owns column i in S[0..15]): the pivot column comes from lane
// J through scalar registers, read entry by entry before lane J overwrites
// it; the owner starts from the unit vector e_J (its base is multiplied by
// 0), so every lane runs the same arithmetic:
//   a_JJ <- 1 / a_JJ, a_Ji <- a_Ji / a_JJ, a_KJ <- -a_KJ / a_JJ,
//   a_Ki <- a_Ki - a_KJ a_Ji / a_JJ.
template <typename T, int J, int... K>
__device__ __forceinline__ void gj16_step(T (&S)[16], int i, std::integer_sequence<int, K...>) {
    const bool own = i == J;
    const T keep = own ? T(0) : T(1);
    const T rt = read_lane(S[J], J);
    const T pv = (own ? T(1) : S[J]) / rt;
    ((K != J ? (void)(S[K] = S[K] * keep + -read_lane(S[K], J) * pv) : (void)0), ...);
    S[J] = pv;
}
template <typename T, int... J>
__device__ __forceinline__ void gj16(T (&S)[16], int i, std::integer_sequence<int, J...>) {
    (gj16_step<T, J>(S, i, std::make_integer_sequence<int, 16>{}), ...);
}

// The 48 x 48 (16-vertex bank) elimination with the nine tiles held in
// registers in accumulator layout; LDS only stages the tiles that serve as
// MFMA operands (the pivot inverse, the row tiles A_kt, the column tiles
// A_tk and the new row tiles), 10.9 KB per wavefront for FP64.  The pivot
// tile is inverted column per lane (16 lanes) by scalar Gauss-Jordan with the
// pivot column broadcast through scalar registers.
template <typename T>
__global__ void __launch_bounds__(64) k_bank_inverse_mfma48(int nbanks, const double* __restrict__ bank_upper,
                                                            double* __restrict__ inv64, float* __restrict__ inv32,
                                                            unsigned short* __restrict__ inv16,
                                                            float* __restrict__ inv16_d) {
    typedef typename Mf<T>::acc acc;
    constexpr int D = 48, B = 16, npack = B * (B + 1) / 2, NP = D * (D + 1) / 2, LDW = 17;
    __shared__ T sw[5][16 * LDW];  // pivot inverse, row tiles (2), column tiles (2)
    __shared__ double dg[D];       // half storage: sqrt of the inverse's diagonal
    const Tiles<T, LDW> W{sw[0]}, R0{sw[1]}, R1{sw[2]}, C0{sw[3]}, C1{sw[4]};
    const int b = blockIdx.x, l = threadIdx.x, q = l >> 4, j = l & 15;
    const double* src = bank_upper + (size_t)b * npack * 9;
    acc A[3][3];
#pragma unroll
    for (int ti = 0; ti < 3; ++ti)
#pragma unroll
        for (int tj = 0; tj < 3; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) A[ti][tj][r] = (T)bank_elem<D>(src, 16 * ti + Mf<T>::row(q, r), 16 * tj + j);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int t0 = (k + 1) % 3, t1 = (k + 2) % 3;
        // pivot tile: accumulator layout -> LDS -> column per lane -> inverse
        W.store_c(0, 0, l, A[k][k]);
        __syncthreads();
        T col[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) col[i] = l < 16 ? W.at(i, l) : T(0);
        __syncthreads();
        gj16<T>(col, l, std::make_integer_sequence<int, 16>{});
        if (l < 16) {
#pragma unroll
            for (int i = 0; i < 16; ++i) W.at(i, l) = col[i];
        }
        R0.store_c(0, 0, l, A[k][t0]);
        R1.store_c(0, 0, l, A[k][t1]);
        C0.store_c(0, 0, l, A[t0][k]);
        C1.store_c(0, 0, l, A[t1][k]);
        __syncthreads();
        T pa[4], pb[4], rb0[4], rb1[4], ca0[4], ca1[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            pa[c] = W.opa(0, 0, c, l);
            pb[c] = W.opb(0, 0, c, l);
            rb0[c] = R0.opb(0, 0, c, l);
            rb1[c] = R1.opb(0, 0, c, l);
            ca0[c] = C0.opa(0, 0, c, l);
            ca1[c] = C1.opa(0, 0, c, l);
        }
        const acc P = W.load_c(0, 0, l);
        acc Tr0 = {0, 0, 0, 0}, Tr1 = {0, 0, 0, 0}, Tc0 = {0, 0, 0, 0}, Tc1 = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            Tr0 = Mf<T>::mac(pa[c], rb0[c], Tr0);  // P^-1 A_k,t0
            Tr1 = Mf<T>::mac(pa[c], rb1[c], Tr1);
            Tc0 = Mf<T>::mac(ca0[c], pb[c], Tc0);  // A_t0,k P^-1 (old column tile)
            Tc1 = Mf<T>::mac(ca1[c], pb[c], Tc1);
        }
        __syncthreads();
        R0.store_c(0, 0, l, Tr0);
        R1.store_c(0, 0, l, Tr1);
        __syncthreads();
        T trb0[4], trb1[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            trb0[c] = R0.opb(0, 0, c, l);
            trb1[c] = R1.opb(0, 0, c, l);
        }
        // trailing update with the old column tiles and the new row tiles
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            A[t0][t0] = Mf<T>::mac(-ca0[c], trb0[c], A[t0][t0]);
            A[t0][t1] = Mf<T>::mac(-ca0[c], trb1[c], A[t0][t1]);
            A[t1][t0] = Mf<T>::mac(-ca1[c], trb0[c], A[t1][t0]);
            A[t1][t1] = Mf<T>::mac(-ca1[c], trb1[c], A[t1][t1]);
        }
        A[k][t0] = Tr0;
        A[k][t1] = Tr1;
        A[t0][k] = -Tc0;
        A[t1][k] = -Tc1;
        A[k][k] = P;
        __syncthreads();
    }
    // upper triangle (row <= column) packed by rows
    double* o64 = inv64 + (size_t)b * NP;
    float* o32 = inv32 + (size_t)b * NP;
#pragma unroll
    for (int ti = 0; ti < 3; ++ti)
#pragma unroll
        for (int tj = ti; tj < 3; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ti + Mf<T>::row(q, r), c = 16 * tj + j;
                if (row <= c) {
                    const int e = row * (2 * D - row + 1) / 2 + (c - row);
                    if constexpr (sizeof(T) == 8) o64[e] = A[ti][tj][r];
                    o32[e] = (float)A[ti][tj][r];
                }
            }
    if (inv16 == nullptr) return;
    // the half-storage form D S D of the same inverse (the arithmetic of
    // k_pack_inv_half: D = sqrt of the diagonal, 1 where it is not positive;
    // S = D^-1 inv D^-1 rounded to FP16 within its range)
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * t + Mf<T>::row(q, r);
            if (row == 16 * t + j) {
                const double a = (double)A[t][t][r];
                const double di = a > 0.0 ? sqrt(a) : 1.0;
                dg[row] = di;
                inv16_d[(size_t)b * D + row] = (float)di;
            }
        }
    __syncthreads();
    unsigned short* o16 = inv16 + (size_t)b * NP;
#pragma unroll
    for (int ti = 0; ti < 3; ++ti)
#pragma unroll
        for (int tj = ti; tj < 3; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ti + Mf<T>::row(q, r), c = 16 * tj + j;
                if (row <= c) {
                    const int e = row * (2 * D - row + 1) / 2 + (c - row);
                    const float v = (float)((double)A[ti][tj][r] / (dg[row] * dg[c]));
                    const _Float16 hv = (_Float16)fminf(fmaxf(v, -65504.f), 65504.f);
                    o16[e] = __builtin_bit_cast(unsigned short, hv);
                }
            }
}

// One refinement step of the FP32 inverse X0 (inv32, packed upper):
// R = I - A X0 in FP64 MFMA (A the FP64 bank matrix, X0 exact in FP64),
// X1 = X0 + X0 R (upper tiles, FP64 MFMA), stored back to inv32.
template <int NT>
__global__ void __launch_bounds__(64) k_bank_refine(int nbanks, const double* __restrict__ bank_upper,
                                                    float* __restrict__ inv32) {
    typedef typename Mf<double>::acc acc;
    constexpr int D = 16 * NT, B = D / 3, npack = B * (B + 1) / 2, NP = D * (D + 1) / 2;
    constexpr int LD = D + 1;
    __shared__ double sa[D * LD];  // A, then R row tile by row tile
    __shared__ double sx[D * LD];  // X0
    const Tiles<double, LD> A{sa}, X{sx};
    const int b = blockIdx.x, l = threadIdx.x;
    const double* src = bank_upper + (size_t)b * npack * 9;
    float* out = inv32 + (size_t)b * NP;
    for (int e = l; e < D * D; e += 64) {
        const int i = e / D, j = e % D;
        A.at(i, j) = bank_elem<D>(src, i, j);
    }
    for (int e = l; e < NP; e += 64) {
        const int r = packed_row<D>(e);
        const int c = r + (e - r * (2 * D - r + 1) / 2);
        const double v = (double)out[e];
        X.at(r, c) = v;
        X.at(c, r) = v;
    }
    __syncthreads();
    // R row tile i only reads A's row tile i, so it replaces it in place
    const int q = l >> 4, jj = l & 15;
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        acc R[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            acc a;
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = (i == t && Mf<double>::row(q, r) == jj) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < NT; ++k)
#pragma unroll
                for (int ch = 0; ch < 4; ++ch) a = Mf<double>::mac(-A.opa(i, k, ch, l), X.opb(k, t, ch, l), a);
            R[t] = a;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NT; ++t) A.store_c(i, t, l, R[t]);
        __syncthreads();
    }
    // X1 = X0 + X0 R on the upper tiles, staged in A's rows as they finish
    // (row tile i of X1 needs X0's row tile i and all of R: X0 is unchanged,
    // R is read from A, so the results wait in registers until all are done)
#pragma unroll
    for (int i = 0; i < NT; ++i) {
#pragma unroll
        for (int t = i; t < NT; ++t) {
            acc a = X.load_c(i, t, l);
#pragma unroll
            for (int k = 0; k < NT; ++k)
#pragma unroll
                for (int ch = 0; ch < 4; ++ch) a = Mf<double>::mac(X.opa(i, k, ch, l), A.opb(k, t, ch, l), a);
            // tile (i, t) of X1 goes to the output directly (upper elements)
            const int c = 16 * t + jj;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * i + Mf<double>::row(q, r);
                if (row <= c) out[row * (2 * D - row + 1) / 2 + (c - row)] = (float)a[r];
            }
        }
    }
}

}  // namespace

void bank_inverse_mfma(Ctx& c, int bank, int nbanks, const double* bank_upper, double* inv64, float* inv32,
                       int mode, unsigned short* inv16, float* inv16_d) {
    if (nbanks <= 0) return;
    hipStream_t s = c.stream;
    if (bank != 16 && bank != 32) throw arg_error("bank inverse: bank size must be 16 or 32");
    if (mode == BANK_INV_MFMA64) {
        if (bank == 16)
            hipLaunchKernelGGL((k_bank_inverse_mfma48<double>), dim3(nbanks), dim3(64), 0, s, nbanks, bank_upper,
                               inv64, inv32, inv16, inv16_d);
        else
            hipLaunchKernelGGL((k_bank_inverse_mfma<double, 6>), dim3(nbanks), dim3(64), 0, s, nbanks, bank_upper,
                               inv64, inv32);
        SGIPC_LAUNCH_CHECK();
    } else if (mode == BANK_INV_MFMA32_REFINED) {
        if (bank == 16) {
            hipLaunchKernelGGL((k_bank_inverse_mfma48<float>), dim3(nbanks), dim3(64), 0, s, nbanks, bank_upper,
                               inv64, inv32, (unsigned short*)nullptr, (float*)nullptr);  // refined below
            SGIPC_LAUNCH_CHECK();
            hipLaunchKernelGGL((k_bank_refine<3>), dim3(nbanks), dim3(64), 0, s, nbanks, bank_upper, inv32);
        } else {
            hipLaunchKernelGGL((k_bank_inverse_mfma<float, 6>), dim3(nbanks), dim3(64), 0, s, nbanks, bank_upper,
                               inv64, inv32);
            SGIPC_LAUNCH_CHECK();
            hipLaunchKernelGGL((k_bank_refine<6>), dim3(nbanks), dim3(64), 0, s, nbanks, bank_upper, inv32);
        }
        SGIPC_LAUNCH_CHECK();
    } else {
        throw arg_error("bank inverse: unknown MFMA mode");
    }
}

}  // namespace sgipc
