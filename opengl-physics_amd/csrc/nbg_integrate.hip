// nbg_integrate.hip -- leapfrog kick/drift kernels and FP64 diagnostics for gfx950.
//
// LeapfrogStep (projects/nbody_core/src/Integrator.cpp:9-39) is kick(dt/2), drift(dt),
// a(x), kick(dt/2).  The first kick and the drift touch the same particle only, so they
// run as one HBM pass (bit-identical per particle to running them as two loops).
// Diagnostics follow System::TotalEnergy / CenterOfMass / AngularMomentum
// (projects/nbody_core/include/ngrav/System.hpp:119-193) in FP64.
#include <cmath>
#include <vector>

#include "nbg_common.hpp"
#include "nbg_internal.hpp"

namespace nbg {

__global__ __launch_bounds__(256) void k_kick_drift(float* __restrict__ x, float* __restrict__ y,
                                                    float* __restrict__ z, float* __restrict__ vx,
                                                    float* __restrict__ vy, float* __restrict__ vz,
                                                    const float* __restrict__ ax, const float* __restrict__ ay,
                                                    const float* __restrict__ az, int n, float hdt, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float nvx = vx[i] + hdt * ax[i];
    const float nvy = vy[i] + hdt * ay[i];
    const float nvz = vz[i] + hdt * az[i];
    vx[i] = nvx;
    vy[i] = nvy;
    vz[i] = nvz;
    x[i] = x[i] + dt * nvx;
    y[i] = y[i] + dt * nvy;
    z[i] = z[i] + dt * nvz;
}

__global__ __launch_bounds__(256) void k_kick(float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
                                              const float* __restrict__ ax, const float* __restrict__ ay,
                                              const float* __restrict__ az, int n, float hdt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] = vx[i] + hdt * ax[i];
    vy[i] = vy[i] + hdt * ay[i];
    vz[i] = vz[i] + hdt * az[i];
}

__global__ __launch_bounds__(256) void k_copy_f32(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void launch_kick(nbg_ctx* c, float hdt) {
    hipLaunchKernelGGL(k_kick, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->vx, c->vy, c->vz, c->ax, c->ay,
                       c->az, c->n, hdt);
}

void launch_kick_drift(nbg_ctx* c, float hdt, float dt) {
    hipLaunchKernelGGL(k_kick_drift, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->x, c->y, c->z, c->vx,
                       c->vy, c->vz, c->ax, c->ay, c->az, c->n, hdt, dt);
}

void launch_copy_f32(nbg_ctx* c, float* dst, const float* src, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_f32, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, dst, src, n);
}

// ---------------------------------------------------------------------------- diagnostics
constexpr int kDiagThreads = 256;
constexpr int kDiagVals = 14;  // M, Mx,My,Mz, Px,Py,Pz, L0x,L0y,L0z, KE, PE, pad, pad

__device__ double block_sum_d(double v, double* s_red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) s_red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kDiagThreads / 64; ++w) t += s_red[w];
    __syncthreads();
    return t;
}

// Per-block FP64 partials of mass, first moments, momentum, angular momentum about the
// origin and kinetic energy, over particles [lo, hi).
__global__ __launch_bounds__(kDiagThreads) void k_diag_moments(const float* __restrict__ x, const float* __restrict__ y,
                                                               const float* __restrict__ z, const float* __restrict__ vx,
                                                               const float* __restrict__ vy,
                                                               const float* __restrict__ vz,
                                                               const float* __restrict__ m, int lo, int hi,
                                                               double* __restrict__ part) {
    __shared__ double s_red[kDiagThreads / 64];
    double v[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
        const double mi = m[i], px = x[i], py = y[i], pz = z[i], ux = vx[i], uy = vy[i], uz = vz[i];
        v[0] += mi;
        v[1] += mi * px;
        v[2] += mi * py;
        v[3] += mi * pz;
        v[4] += mi * ux;
        v[5] += mi * uy;
        v[6] += mi * uz;
        v[7] += mi * (py * uz - pz * uy);
        v[8] += mi * (pz * ux - px * uz);
        v[9] += mi * (px * uy - py * ux);
        v[10] += 0.5 * mi * (ux * ux + uy * uy + uz * uz);
    }
    for (int k = 0; k < 11; ++k) {
        const double t = block_sum_d(v[k], s_red);
        if (threadIdx.x == 0) part[blockIdx.x * kDiagVals + k] = t;
    }
}

// phi(q) h / (G m) of the cubic-spline kernel, SplinePotential3D
// (projects/nbody_core/include/ngrav/Softening.hpp:49-56).
__device__ __forceinline__ double spline_phi(double q) {
    if (q >= 2.0) return -1.0 / q;
    if (q < 1.0) return -(1.4 - (2.0 / 3.0) * q * q + 0.3 * q * q * q * q - 0.1 * q * q * q * q * q);
    const double t = 2.0 - q;
    const double t2 = t * t, t4 = t2 * t2;
    const double F = t4 / 4.0 - t4 * t / 5.0 + t4 * t2 / 24.0;
    return -((1.0 - 4.0 * F) / q + t4 / 2.0 - t4 * t / 5.0);
}

// Pairwise potential, all ordered pairs i != j (each unordered pair twice, halved on the
// host), FP64: Plummer -G m_i m_j / sqrt(r^2 + eps2); spline G m_i m_j phi(r / h) / h
// (System::TotalEnergy, System.hpp:141-162).  Sources stream through scalar loads.
template <bool SPL>
__global__ __launch_bounds__(kDiagThreads) void k_diag_potential(const float* __restrict__ x,
                                                                 const float* __restrict__ y,
                                                                 const float* __restrict__ z,
                                                                 const float* __restrict__ m, int lo, int hi,
                                                                 double eps2, double h, double* __restrict__ part) {
    __shared__ double s_red[kDiagThreads / 64];
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < hi;
    const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, zi = valid ? z[i] : 0.0;
    const double mi = valid ? m[i] : 0.0;
    double acc = 0.0;
    for (int j = lo; j < hi; ++j) {
        const double dx = static_cast<double>(x[j]) - xi;
        const double dy = static_cast<double>(y[j]) - yi;
        const double dz = static_cast<double>(z[j]) - zi;
        double term;
        if (SPL) {
            const double r = __dsqrt_rn(dx * dx + dy * dy + dz * dz);
            const double q = (r <= 1e-300) ? 0.0 : r / h;
            term = static_cast<double>(m[j]) * spline_phi(q);
        } else {
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            // 1/sqrt(r2) to FP64 accuracy: hardware estimate + two Newton steps
            double r = __builtin_amdgcn_rsq(r2);
            r = r * (1.5 - 0.5 * r2 * r * r);
            r = r * (1.5 - 0.5 * r2 * r * r);
            term = -static_cast<double>(m[j]) * r;
        }
        acc += (j == i) ? 0.0 : term;
    }
    acc *= SPL ? mi / h : mi;
    const double t = block_sum_d(acc, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

int64_t diag_partials_needed(int64_t n) {
    return std::max<int64_t>(1024 * kDiagVals, grid_for(n, kDiagThreads) + 16);
}

int diagnostics(nbg_ctx* c, const nbg_params& p, int with_potential, double* out, int sys) {
    const int lo = static_cast<int>(c->h_sys_off[static_cast<size_t>(sys)]);
    const int hi = static_cast<int>(c->h_sys_off[static_cast<size_t>(sys) + 1]);
    const int nb = 1024;
    hipLaunchKernelGGL(k_diag_moments, dim3(nb), dim3(kDiagThreads), 0, c->stream, c->x, c->y, c->z, c->vx, c->vy,
                       c->vz, c->m, lo, hi, c->diag_part);
    std::vector<double> h(static_cast<size_t>(nb) * kDiagVals);
    NBG_HIP_CHECK(c, hipMemcpyAsync(h.data(), c->diag_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    double s[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 11; ++k) s[k] += h[static_cast<size_t>(b) * kDiagVals + k];
    const double M = s[0];
    double com[3] = {s[1], s[2], s[3]};
    if (M > 0.0)
        for (int k = 0; k < 3; ++k) com[k] /= M;
    const double P[3] = {s[4], s[5], s[6]};
    // L about the COM = L about the origin - com x P
    const double L[3] = {s[7] - (com[1] * P[2] - com[2] * P[1]), s[8] - (com[2] * P[0] - com[0] * P[2]),
                         s[9] - (com[0] * P[1] - com[1] * P[0])};
    double pot = 0.0;
    if (with_potential && hi > lo) {
        const int npb = grid_for(hi - lo, kDiagThreads);
        if (p.softening == NBG_SOFT_SPLINE)
            hipLaunchKernelGGL(k_diag_potential<true>, dim3(npb), dim3(kDiagThreads), 0, c->stream, c->x, c->y, c->z,
                               c->m, lo, hi, p.eps2, std::sqrt(p.eps2), c->diag_part);
        else
            hipLaunchKernelGGL(k_diag_potential<false>, dim3(npb), dim3(kDiagThreads), 0, c->stream, c->x, c->y,
                               c->z, c->m, lo, hi, p.eps2, 1.0, c->diag_part);
        std::vector<double> hp(static_cast<size_t>(npb));
        NBG_HIP_CHECK(c, hipMemcpyAsync(hp.data(), c->diag_part, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                                        c->stream));
        NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        for (double v : hp) pot += v;
        pot = 0.5 * p.G * pot;
    }
    out[0] = s[10];
    out[1] = pot;
    for (int k = 0; k < 3; ++k) {
        out[2 + k] = com[k];
        out[5 + k] = L[k];
        out[8 + k] = P[k];
    }
    return NBG_OK;
}

}  // namespace nbg
