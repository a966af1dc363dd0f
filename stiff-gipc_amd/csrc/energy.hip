// FP64 energies used by the Newton line search: kinetic and SNK1 elastic
// here, IPC barrier in barrier.hip (sharing its distance code); each a
// fixed-order reduction (per-workgroup partials summed
// by one workgroup in index order), so repeated evaluations are bitwise
// identical.
//
// Reference: GIPC::computeEnergy (S/GIPC.cu:10520-10579) over
// Energy_Add_Reduction_Algorithm (:10347-10470): type 0
// _getKineticEnergy_Reduction_3D (:7330-7360), type 1
// _getFEMEnergy_Reduction_3D (:7515-7583) with __cal_StabbleNHK_energy1_3D
// (S/femEnergy.cu:1032-1050), scaled by dt^2 after the sum, type 2
// _getBarrierEnergy_Reduction_3D (:7768-7800) with __cal_Barrier_energy
// (:701-957, RANK 2).
#include "context.hpp"
#include "pcg_dev.hpp"
#include <mutex>

namespace sgipc {
namespace {

constexpr int kEThreads = 256;
constexpr int kEParts = 1024;  // partials per energy

struct E3 {
    double x, y, z;
};
__device__ inline E3 ld3(const double* p, int i) { return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]}; }
__device__ inline E3 sub(E3 a, E3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline double dot(E3 a, E3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// SNK1: 0.5 vol [mu (I2 - 3) + lam (J - 1 - mu/lam)^2], F = Ds DmInv
__global__ void __launch_bounds__(kEThreads) k_fem_energy(int T, const double* __restrict__ x,
                                                          const uint4* __restrict__ tets,
                                                          const double* __restrict__ DmInv,
                                                          const double* __restrict__ vol, const double* __restrict__ mu,
                                                          const double* __restrict__ lam, double* __restrict__ part) {
    __shared__ double sh[kEThreads / 64];
    double e = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        uint4 q = tets[t];
        E3 x0 = ld3(x, q.x), x1 = ld3(x, q.y), x2 = ld3(x, q.z), x3 = ld3(x, q.w);
        E3 c0 = sub(x1, x0), c1 = sub(x2, x0), c2 = sub(x3, x0);  // Ds columns
        const double* D = DmInv + 9 * (size_t)t;                 // row-major
        double F[9];  // row-major F(i, j) = sum_k Ds(i, k) D(k, j)
        const double ds[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) F[3 * i + j] = ds[i][0] * D[j] + ds[i][1] * D[3 + j] + ds[i][2] * D[6 + j];
        double I2 = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) I2 += F[i] * F[i];
        double J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                   F[2] * (F[3] * F[7] - F[4] * F[6]);
        double m = mu[t], l = lam[t], Jm = J - 1.0 - m / l;
        e += 0.5 * (m * (I2 - 3.0) + l * Jm * Jm) * vol[t];
    }
    write_partial(e, part, sh);
}

__global__ void __launch_bounds__(kEThreads) k_kinetic_energy(int n, const double* __restrict__ x,
                                                              const double* __restrict__ xt,
                                                              const double* __restrict__ m, double* __restrict__ part) {
    __shared__ double sh[kEThreads / 64];
    double e = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        E3 d = sub(ld3(x, i), ld3(xt, i));
        e += dot(d, d) * m[i] * 0.5;
    }
    write_partial(e, part, sh);
}

__global__ void __launch_bounds__(kEThreads) k_energy_total(const double* __restrict__ part, int np,
                                                            double* __restrict__ out) {
    __shared__ double sh[kEThreads / 64];
    double s = sum_partials(part, np, sh);
    if (threadIdx.x == 0) *out = s;
}

// fixed-order sum of the partials, times `scale`, to device memory
__global__ void __launch_bounds__(kEThreads) k_energy_total_scaled(const double* __restrict__ part, int np,
                                                                   double scale, double* __restrict__ out) {
    __shared__ double sh[kEThreads / 64];
    double s = sum_partials(part, np, sh);
    if (threadIdx.x == 0) *out = scale * s;
}

// Scratch of the allocation-free FP64 SNK1 energy: one partial per
// workgroup (at most kEParts).  Each device holds its own copy at its own
// address, looked up once per device by energy_scratch_setup.
static __device__ __align__(256) double scratch[kEParts];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Device address of this device's energy scratch (device current).
void energy_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw arg_error("energy scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    SGIPC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw hip_error("energy scratch: no device address");
    g_scratch_ptr[device] = static_cast<double*>(p);
}

// dt^2 sum_t SNK1 energy to device memory, partials in the static scratch:
// no allocation, no synchronisation.  Same grid and reduction order as
// fem_energy, so the same bits.
void fem_energy_noalloc(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                        const double* mu, const double* lam, double dt, double* out) {
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c.device >= 0 && c.device < kMaxDevices) part = g_scratch_ptr[c.device];
    }
    if (!part) throw state_error("fem_energy_noalloc: scratch not set up on this device");
    const int g = energy_grid(T);
    if (g > kEParts) throw capacity_error("fem_energy_noalloc: more partials than the scratch holds");
    hipLaunchKernelGGL(k_fem_energy, dim3(g), dim3(kEThreads), 0, c.stream, T, x, (const uint4*)tets, DmInv, vol, mu,
                       lam, part);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_energy_total_scaled, dim3(1), dim3(kEThreads), 0, c.stream, (const double*)part, g, dt * dt,
                       out);
    SGIPC_LAUNCH_CHECK();
}

// ---- line-search step (no energies) ----------------------------------------
// x = x_temp - alpha dir on the free vertices (|btype| == 0), on every vertex
// when move_boundary (_stepForward, S/GIPC.cu:7937-7954)
__global__ void k_step_forward(int n, double* __restrict__ x, const double* __restrict__ xt,
                               const double* __restrict__ dir, const int* __restrict__ btype, double alpha,
                               int move_boundary) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (!move_boundary && btype && btype[v] != 0) return;
    const size_t o = 3 * (size_t)v;
    x[o] = xt[o] - alpha * dir[o];
    x[o + 1] = xt[o + 1] - alpha * dir[o + 1];
    x[o + 2] = xt[o + 2] - alpha * dir[o + 2];
}

// per-workgroup max of |dir_v| over the listed vertices (the CFL speed of
// cfl_largestSpeed, S/GIPC.cu:8916-8947; _reduct_max_cfl_to_double :7157)
__global__ void __launch_bounds__(kEThreads) k_max_speed(int m, const uint32_t* __restrict__ ids,
                                                          const double* __restrict__ dir, double* __restrict__ part) {
    __shared__ double sh[kEThreads / 64];
    double v = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const size_t o = 3 * (size_t)ids[i];
        const double a = dir[o], b = dir[o + 1], c = dir[o + 2];
        v = fmax(v, sqrt(a * a + b * b + c * c));
    }
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double w = sh[0];
        for (int i = 1; i < kEThreads / 64; ++i) w = fmax(w, sh[i]);
        part[blockIdx.x] = w;
    }
}

__global__ void __launch_bounds__(kEThreads) k_max_total(const double* __restrict__ part, int g,
                                                          double* __restrict__ out) {
    __shared__ double sh[kEThreads / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < g; i += blockDim.x) v = fmax(v, part[i]);
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double w = sh[0];
        for (int i = 1; i < kEThreads / 64; ++i) w = fmax(w, sh[i]);
        *out = w;
    }
}

int energy_grid(long long n) {
    long long g = div_up(n, kEThreads);
    return (int)(g < 1 ? 1 : (g > kEParts ? kEParts : g));
}

double* energy_partials(Ctx& c) { return c.energy_part.reserve(kEParts + 1); }

// sum of g per-workgroup partials (fixed order) to the host
double energy_finish(Ctx& c, double* part, int g) {
    double* out = part + kEParts;
    hipLaunchKernelGGL(k_energy_total, dim3(1), dim3(kEThreads), 0, c.stream, (const double*)part, g, out);
    SGIPC_LAUNCH_CHECK();
    double h = 0.0;
    SGIPC_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    return h;
}

double kinetic_energy(Ctx& c, int n, const double* x, const double* xt, const double* m) {
    double* part = energy_partials(c);
    int g = energy_grid(n);
    hipLaunchKernelGGL(k_kinetic_energy, dim3(g), dim3(kEThreads), 0, c.stream, n, x, xt, m, part);
    SGIPC_LAUNCH_CHECK();
    return energy_finish(c, part, g);
}

double fem_energy(Ctx& c, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                  const double* mu, const double* lam, double dt) {
    double* part = energy_partials(c);
    int g = energy_grid(T);
    hipLaunchKernelGGL(k_fem_energy, dim3(g), dim3(kEThreads), 0, c.stream, T, x, (const uint4*)tets, DmInv, vol, mu,
                       lam, part);
    SGIPC_LAUNCH_CHECK();
    return dt * dt * energy_finish(c, part, g);
}

void step_forward(Ctx& c, int n, double* x, const double* xt, const double* dir, const int* btype, double alpha,
                  int move_boundary) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_step_forward, dim3(div_up(n, 256)), dim3(256), 0, c.stream, n, x, xt, dir, btype, alpha,
                       move_boundary);
    SGIPC_LAUNCH_CHECK();
}

double max_speed(Ctx& c, int m, const uint32_t* ids, const double* dir) {
    if (m <= 0) return 0.0;
    double* part = energy_partials(c);
    const int g = energy_grid(m);
    hipLaunchKernelGGL(k_max_speed, dim3(g), dim3(kEThreads), 0, c.stream, m, ids, dir, part);
    SGIPC_LAUNCH_CHECK();
    double* out = part + kEParts;
    hipLaunchKernelGGL(k_max_total, dim3(1), dim3(kEThreads), 0, c.stream, (const double*)part, g, out);
    SGIPC_LAUNCH_CHECK();
    double h = 0.0;
    SGIPC_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    return h;
}

}  // namespace sgipc
