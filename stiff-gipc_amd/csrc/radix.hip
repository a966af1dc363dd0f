// Stable device radix sort of (key, 32-bit value) pairs: least-significant
// digit first, 8-bit digits, one launch per digit.
//
// Launches: a fill of the pass state, one digit histogram of every pass in
// a single read of the keys, one exclusive scan of those histograms (the
// first output position of every digit), then per pass one "onesweep"
// kernel (Adinets and Merrill, "Onesweep: A Faster Least Significant Digit
// Radix Sort for GPUs", 2022): each workgroup takes the next tile of 4096
// keys from a counter, ranks its keys by digit (stable: wave-striped rounds
// of 64 consecutive keys, equal digits matched with eight ballots), publishes
// the tile's digit counts and looks back over the earlier tiles' published
// counts for its global offsets (decoupled look-back), then stages the tile
// in LDS in digit order and writes it out in coalesced runs.
//
// The look-back waits only on tiles taken earlier, by workgroups already
// running, which publish before they look back themselves, so every wait
// ends; the spin is still bounded, and an exhausted bound sets an error word
// instead of waiting on (the sort is then wrong and the caller throws).
// Status words are agent-scope relaxed atomics: flag in bits 31:30, count in
// bits 29:0 (n < 2^30).
#include "context.hpp"
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace sgipc {
namespace {

constexpr int kRT = 256;               // threads per workgroup (one per digit)
constexpr int kRW = kRT / 64;          // wavefronts
constexpr int kRI = 16;                // keys per thread
constexpr int kRTile = kRT * kRI;      // 4096 keys per tile
constexpr unsigned kAgg = 1u << 30;    // tile count published
constexpr unsigned kPre = 2u << 30;    // inclusive prefix through this tile published
constexpr unsigned kVal = (1u << 30) - 1u;
constexpr long long kMaxSpin = 1LL << 24;

// digit of pass p: 8 bits, fewer in the last pass (keys sort by bits [0, bits))
__device__ __host__ inline unsigned digit_mask(int bits, int p) {
    const int b = bits - 8 * p;
    return b >= 8 ? 255u : ((1u << b) - 1u);
}

template <typename K>
__global__ void __launch_bounds__(256) k_digit_hist(const K* __restrict__ kin, long long n, int bits,
                                                     int* __restrict__ hist) {
    const int passes = (bits + 7) / 8;
    __shared__ int h[8 * 256];
    for (int i = threadIdx.x; i < passes * 256; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const K k = kin[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p * 256 + (int)((k >> (8 * p)) & digit_mask(bits, p))], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * 256; i += blockDim.x)
        if (h[i]) atomicAdd(hist + i, h[i]);
}

// exclusive scan of 256 ints by 256 threads (fixed order)
__device__ inline int scan256(int v, int* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < w; ++q) before += sh[q];
    __syncthreads();
    return before + incl - v;
}

// first output position of every digit of every pass (one workgroup per pass)
__global__ void __launch_bounds__(256) k_digit_starts(const int* __restrict__ hist, int* __restrict__ dstart) {
    __shared__ int sh[4];
    const int p = blockIdx.x, d = threadIdx.x;
    dstart[p * 256 + d] = scan256(hist[p * 256 + d], sh);
}

template <typename K>
__global__ void __launch_bounds__(kRT) k_onesweep(const K* __restrict__ kin, const int* __restrict__ vin, int id_base,
                                                  K* __restrict__ kout, int* __restrict__ vout, long long n, int shift,
                                                  unsigned dmask, unsigned* __restrict__ status, const int* __restrict__ dstart,
                                                  int* __restrict__ tile_ctr, int* __restrict__ err) {
    __shared__ int cnt[kRW][256];  // per wavefront and digit: keys so far, then the wavefront's offset
    __shared__ int toff[256];      // first tile position of each digit
    __shared__ long long gofs[256];
    __shared__ int sh[4];
    __shared__ K sk[kRTile];
    __shared__ int sv[kRTile];
    __shared__ int tile_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) tile_s = atomicAdd(tile_ctr, 1);
    for (int i = lane; i < 256; i += 64) cnt[w][i] = 0;
    __syncthreads();
    const int tile = tile_s;
    const long long base = (long long)tile * kRTile + (long long)w * (kRTile / kRW);
    const unsigned long long lt = (1ull << lane) - 1ull;
    K kr[kRI];
    int vr[kRI], rk[kRI];
#pragma unroll
    for (int j = 0; j < kRI; ++j) {
        const long long idx = base + j * 64 + lane;
        const bool valid = idx < n;
        kr[j] = valid ? kin[idx] : K(0);
        vr[j] = valid ? (vin ? vin[idx] : id_base + (int)idx) : 0;
    }
    // stable ranks: rounds in key order, lanes in key order within a round
#pragma unroll
    for (int j = 0; j < kRI; ++j) {
        const bool valid = base + j * 64 + lane < n;
        const int d = (int)((kr[j] >> shift) & dmask);
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        int r = 0;
        if (valid) {
            const int before = __popcll(m & lt);
            const int c0 = cnt[w][d];
            r = c0 + before;
            if (before == 0) cnt[w][d] = c0 + __popcll(m);
        }
        rk[j] = r;
    }
    __syncthreads();
    {
        const int d = t;
        int run = 0;
        for (int q = 0; q < kRW; ++q) {
            const int cq = cnt[q][d];
            cnt[q][d] = run;
            run += cq;
        }
        unsigned* mine = status + (size_t)tile * 256 + d;
        long long excl = 0;
        if (tile == 0) {
            __hip_atomic_store(mine, kPre | (unsigned)run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(mine, kAgg | (unsigned)run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            long long spins = 0;
            for (int p = tile - 1; p >= 0;) {
                const unsigned st =
                    __hip_atomic_load(status + (size_t)p * 256 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((st >> 30) == 0u) {
                    if (++spins > kMaxSpin) {
                        atomicOr(err, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (long long)(st & kVal);
                if (st & kPre) break;
                --p;
            }
            __hip_atomic_store(mine, kPre | (unsigned)((excl + run) & kVal), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        gofs[d] = (long long)dstart[d] + excl;
        toff[d] = scan256(run, sh);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRI; ++j) {
        if (base + j * 64 + lane < n) {
            const int d = (int)((kr[j] >> shift) & dmask);
            const int pos = toff[d] + cnt[w][d] + rk[j];
            sk[pos] = kr[j];
            sv[pos] = vr[j];
        }
    }
    __syncthreads();
    const long long left = n - (long long)tile * kRTile;
    const int tn = left < kRTile ? (int)left : kRTile;
    for (int i = t; i < tn; i += kRT) {
        const K key = sk[i];
        const int d = (int)((key >> shift) & dmask);
        const long long dst = gofs[d] + (i - toff[d]);
        kout[dst] = key;
        vout[dst] = sv[i];
    }
}

template <typename K>
void sort_t(Ctx& c, const K* kin, K* kout, const int* vin, int id_base, int* vout, long long n, int bits,
            hipStream_t s, int* err) {
    if (!c.sort_err.p) {  // sticky look-back failure word, cleared once
        c.sort_err.reserve(1);
        fill<int>(c.sort_err.p, 0, 1, s);
    }
    if (!err) err = c.sort_err.p;
    if (n <= 0) return;
    if (n >= (1LL << 30)) throw arg_error("radix_sort: at most 2^30 - 1 pairs");
    if (bits < 1 || bits > 8 * (int)sizeof(K)) throw arg_error("radix_sort: bad key bit count");
    if ((const void*)kin == (const void*)kout) throw arg_error("radix_sort: keys must not be sorted in place");
    const int P = (bits + 7) / 8;
    const long long ntiles = (n + kRTile - 1) / kRTile;
    // scratch: keys and values of the other ping-pong buffer, then the ints
    const size_t kb = ((size_t)n * sizeof(K) + 255) / 256 * 256;
    const size_t vb = ((size_t)n * sizeof(int) + 255) / 256 * 256;
    const long long nstat = (long long)P * ntiles * 256;
    const long long nint = nstat + 2LL * P * 256 + P;
    char* tmp = c.sort_tmp.reserve(kb + vb + (size_t)nint * sizeof(int) + 256);
    K* ktmp = reinterpret_cast<K*>(tmp);
    int* vtmp = reinterpret_cast<int*>(tmp + kb);
    int* ints = reinterpret_cast<int*>(tmp + kb + vb);
    unsigned* status = reinterpret_cast<unsigned*>(ints);
    int* hist = ints + nstat;
    int* dst = hist + P * 256;
    int* ctr = dst + P * 256;
    fill<int>(ints, 0, nint, s);
    const int hg = (int)std::min<long long>(1024, (n + 255) / 256);
    hipLaunchKernelGGL((k_digit_hist<K>), dim3(hg), dim3(256), 0, s, kin, n, bits, hist);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_digit_starts, dim3(P), dim3(256), 0, s, (const int*)hist, dst);
    SGIPC_LAUNCH_CHECK();
    const K* ks = kin;
    const int* vs = vin;
    for (int p = 0; p < P; ++p) {
        const bool to_out = ((P - 1 - p) % 2) == 0;
        K* kd = to_out ? kout : ktmp;
        int* vd = to_out ? vout : vtmp;
        hipLaunchKernelGGL((k_onesweep<K>), dim3((unsigned)ntiles), dim3(kRT), 0, s, ks, vs, id_base, kd, vd, n, 8 * p,
                           digit_mask(bits, p), status + (size_t)p * ntiles * 256, (const int*)(dst + p * 256),
                           ctr + p, err);
        SGIPC_LAUNCH_CHECK();
        ks = kd;
        vs = vd;
    }
}

}  // namespace

void radix_sort_pairs(Ctx& c, const unsigned long long* kin, unsigned long long* kout, const int* vin, int id_base,
                      int* vout, long long n, int bits, hipStream_t s, int* err) {
    sort_t<unsigned long long>(c, kin, kout, vin, id_base, vout, n, bits, s, err);
}
void radix_sort_pairs(Ctx& c, const unsigned* kin, unsigned* kout, const int* vin, int id_base, int* vout,
                      long long n, int bits, hipStream_t s, int* err) {
    sort_t<unsigned>(c, kin, kout, vin, id_base, vout, n, bits, s, err);
}

namespace {
template <typename K>
void rocprim_sort(Ctx& c, const K* kin, K* kout, const int* vin, int* vout, long long n, int bits) {
    hipStream_t s = c.stream;
    size_t bytes = 0;
    K* ki = const_cast<K*>(kin);
    if (vin) {
        int* vi = const_cast<int*>(vin);
        SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, ki, kout, vi, vout, (size_t)n, 0, bits, s));
        char* tmp = c.sort_tmp.reserve(bytes + 1);
        SGIPC_HIP(rocprim::radix_sort_pairs((void*)tmp, bytes, ki, kout, vi, vout, (size_t)n, 0, bits, s));
    } else {
        rocprim::counting_iterator<int> ids(0);
        SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, ki, kout, ids, vout, (size_t)n, 0, bits, s));
        char* tmp = c.sort_tmp.reserve(bytes + 1);
        SGIPC_HIP(rocprim::radix_sort_pairs((void*)tmp, bytes, ki, kout, ids, vout, (size_t)n, 0, bits, s));
    }
}
}  // namespace

void sort_pairs_dispatch(Ctx& c, const void* kin, void* kout, const int* vin, int* vout, long long n, int key_bytes,
                         int bits, int impl) {
    if (n > 0) {
        using U64 = unsigned long long;
        if (impl == SORT_OWN) {
            if (key_bytes == 8) radix_sort_pairs(c, (const U64*)kin, (U64*)kout, vin, 0, vout, n, bits, c.stream);
            else radix_sort_pairs(c, (const unsigned*)kin, (unsigned*)kout, vin, 0, vout, n, bits, c.stream);
        } else {
            if (key_bytes == 8) rocprim_sort<U64>(c, (const U64*)kin, (U64*)kout, vin, vout, n, bits);
            else rocprim_sort<unsigned>(c, (const unsigned*)kin, (unsigned*)kout, vin, vout, n, bits);
        }
    }
    SGIPC_HIP(hipStreamSynchronize(c.stream));
}

int radix_sort_error(Ctx& c) {
    if (!c.sort_err.p) return 0;
    int h = 0;
    SGIPC_HIP(hipMemcpyAsync(&h, c.sort_err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    return h;
}

}  // namespace sgipc
