// Ground (half-space) barrier: contact detection against a plane n.x = o,
// the barrier gradient and 3x3 Hessian block per close vertex, its FP64
// energy for the line search and the largest feasible step toward the plane.
//
// With s = (n.x - o)^2 the squared distance and dhat squared (as
// GIPC::dHat), the energy is kappa * b(s), b(s) = -(s - dhat)^2 ln(s / dhat)
// for s < dhat.  With g_b = db/ds and H_b = d2b/ds2:
//   gradient  kappa * g_b * 2 (n.x - o) * n
//   Hessian   kappa * (4 H_b s + 2 g_b) * n n^T        (not projected)
//
// Reference: _GroundCollisionDetect (StiffGIPC/GIPC.cu:6433-6450),
// _computeGroundGradientAndHessian (:6463-6514), _computeGroundEnergy_Reduction
// (:6837-6900, kappa applied after the sum, :10555),
// _reduct_min_groundTimeStep_to_double (:6902-6960) and
// GIPC::ground_largestFeasibleStepSize (:8986-9030).  Detection here keeps
// the order of the candidate list (scan compaction) where the reference
// appends with an atomic counter, so the triplet order is deterministic.
#include "context.hpp"
#include "pcg_dev.hpp"

namespace sgipc {
namespace {

constexpr int kGThreads = 256;
constexpr int kGParts = 1024;

struct Plane {
    double nx, ny, nz, o;
};

__device__ inline double plane_dist(const Plane& P, const double* __restrict__ x, unsigned v) {
    return P.nx * x[3 * (size_t)v] + P.ny * x[3 * (size_t)v + 1] + P.nz * x[3 * (size_t)v + 2] - P.o;
}

__global__ void k_ground_flag(int m, const unsigned* __restrict__ cand, const double* __restrict__ x, Plane P,
                              double dhat, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double d = plane_dist(P, x, cand[i]);
    flag[i] = (d * d <= dhat) ? 1 : 0;
}

__global__ void k_ground_scatter(int m, const unsigned* __restrict__ cand, const int* __restrict__ flag,
                                 const int* __restrict__ pos, unsigned* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (flag[i]) out[pos[i]] = cand[i];
}

__global__ void k_ground_hessian(int np, const unsigned* __restrict__ gp, const double* __restrict__ x, Plane P,
                                 double dhat, double kappa, const int* __restrict__ btype, int index_offset,
                                 double* __restrict__ vals, int* __restrict__ rows, int* __restrict__ cols,
                                 double* __restrict__ fb) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    unsigned v = gp[i];
    double d = plane_dist(P, x, v);
    double s = d * d;
    double t = s - dhat;
    double lg = log(s / dhat);
    double gb = -2.0 * t * lg - t * t / s;
    double hb = -2.0 * lg - 4.0 * t / s + t * t / (s * s);
    bool pin = btype ? (btype[v] != 0) : false;
    double gs = pin ? 0.0 : kappa * gb * 2.0 * d;
    double hs = pin ? 0.0 : kappa * (4.0 * hb * s + 2.0 * gb);
    const double n[3] = {P.nx, P.ny, P.nz};
    if (!pin) {
        // one term per vertex from this kernel (candidates are distinct)
        atomicAdd(fb + 3 * (size_t)v, gs * n[0]);
        atomicAdd(fb + 3 * (size_t)v + 1, gs * n[1]);
        atomicAdd(fb + 3 * (size_t)v + 2, gs * n[2]);
    }
    double* dst = vals + 9 * (size_t)i;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int p = 0; p < 3; ++p) dst[p + 3 * q] = hs * n[p] * n[q];  // column-major, symmetric
    rows[i] = (int)v + index_offset;
    cols[i] = (int)v + index_offset;
}

__global__ void __launch_bounds__(kGThreads) k_ground_energy(int m, const unsigned* __restrict__ ids,
                                                             const double* __restrict__ x, Plane P, double dhat,
                                                             double* __restrict__ part) {
    __shared__ double sh[kGThreads / 64];
    double e = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        double d = plane_dist(P, x, ids[i]);
        double s = d * d;
        if (s < dhat) {
            double t = s - dhat;
            e += -t * t * log(s / dhat);
        }
    }
    write_partial(e, part, sh);
}

__device__ inline double block_max(double v, double* sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, sh[k]);
    return r;
}

// largest inverse step: max over vertices moving toward the plane along -dir
// of (n.dir) / (slack * dist); 1 when nothing approaches
__global__ void __launch_bounds__(kGThreads) k_ground_inv_step(int m, const unsigned* __restrict__ ids,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ dir, Plane P,
                                                               double slack, double* __restrict__ part) {
    __shared__ double sh[kGThreads / 64];
    double r = 1.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        unsigned v = ids[i];
        double coef = P.nx * dir[3 * (size_t)v] + P.ny * dir[3 * (size_t)v + 1] + P.nz * dir[3 * (size_t)v + 2];
        if (coef > 0.0) r = fmax(r, coef / (plane_dist(P, x, v) * slack));
    }
    double b = block_max(r, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void __launch_bounds__(kGThreads) k_max_total(const double* __restrict__ part, int np,
                                                         double* __restrict__ out) {
    __shared__ double sh[kGThreads / 64];
    double r = 1.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) r = fmax(r, part[i]);
    double b = block_max(r, sh);
    if (threadIdx.x == 0) *out = b;
}

int ground_grid(long long m) {
    long long g = div_up(m, kGThreads);
    return (int)(g < 1 ? 1 : (g > kGParts ? kGParts : g));
}

Plane make_plane(const double* normal, double offset) {
    double l2 = normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2];
    if (!(l2 > 0.0) || fabs(l2 - 1.0) > 1e-12) throw arg_error("ground: the plane normal must have unit length");
    return Plane{normal[0], normal[1], normal[2], offset};
}

}  // namespace

int ground_detect(Ctx& c, int m, const unsigned* cand, const double* x, const double* normal, double offset,
                  double dhat, unsigned* out) {
    Plane P = make_plane(normal, offset);
    if (!(dhat > 0.0)) throw arg_error("ground_detect: dhat must be positive");
    if (m <= 0) return 0;
    int* flag = c.cls_flag.reserve((size_t)m + 1);
    int* pos = c.cls_scan.reserve((size_t)m + 1);
    hipLaunchKernelGGL(k_ground_flag, dim3(div_up(m, 256)), dim3(256), 0, c.stream, m, cand, x, P, dhat, flag);
    SGIPC_LAUNCH_CHECK();
    exclusive_scan(c, flag, pos, m, nullptr, c.stream);
    hipLaunchKernelGGL(k_ground_scatter, dim3(div_up(m, 256)), dim3(256), 0, c.stream, m, cand,
                       (const int*)flag, (const int*)pos, out);
    SGIPC_LAUNCH_CHECK();
    int cnt = 0;
    SGIPC_HIP(hipMemcpyAsync(&cnt, pos + m, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    return cnt;
}

void ground_assemble(Ctx& c, int np, const unsigned* gp, const double* x, const double* normal, double offset,
                     double dhat, double kappa, const int* btype, int index_offset, long long triplet_offset,
                     double* fb) {
    Plane P = make_plane(normal, offset);
    if (!c.t_vals) throw state_error("ground_assemble: no triplet view set");
    if (triplet_offset < 0 || triplet_offset + np > c.t_cap)
        throw capacity_error("ground_assemble: triplet capacity exceeded");
    if (np <= 0) return;
    hipLaunchKernelGGL(k_ground_hessian, dim3(div_up(np, 128)), dim3(128), 0, c.stream, np, gp, x, P, dhat, kappa,
                       btype, index_offset, c.t_vals + 9 * triplet_offset, c.t_rows + triplet_offset,
                       c.t_cols + triplet_offset, fb);
    SGIPC_LAUNCH_CHECK();
}

double ground_energy(Ctx& c, int m, const unsigned* ids, const double* x, const double* normal, double offset,
                     double dhat, double kappa) {
    Plane P = make_plane(normal, offset);
    double* part = energy_partials(c);
    int g = ground_grid(m);
    hipLaunchKernelGGL(k_ground_energy, dim3(g), dim3(kGThreads), 0, c.stream, m, ids, x, P, dhat, part);
    SGIPC_LAUNCH_CHECK();
    return kappa * energy_finish(c, part, g);
}

double ground_step_bound(Ctx& c, int m, const unsigned* ids, const double* x, const double* dir,
                         const double* normal, double offset, double slack) {
    Plane P = make_plane(normal, offset);
    if (!(slack > 0.0 && slack <= 1.0)) throw arg_error("ground_step_bound: slack must be in (0, 1]");
    if (m <= 0) return 1.0;
    double* part = energy_partials(c);
    int g = ground_grid(m);
    hipLaunchKernelGGL(k_ground_inv_step, dim3(g), dim3(kGThreads), 0, c.stream, m, ids, x, dir, P, slack, part);
    SGIPC_LAUNCH_CHECK();
    double* out = part + kGParts;
    hipLaunchKernelGGL(k_max_total, dim3(1), dim3(kGThreads), 0, c.stream, (const double*)part, g, out);
    SGIPC_LAUNCH_CHECK();
    double h = 1.0;
    SGIPC_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SGIPC_HIP(hipStreamSynchronize(c.stream));
    return 1.0 / h;
}

}  // namespace sgipc
