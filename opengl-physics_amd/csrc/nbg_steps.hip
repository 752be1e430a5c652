// nbg_steps.hip -- adaptive global time steps and block (rung) time steps for gfx950.
//
//   ChooseAdaptiveDt  projects/nbody_core/src/Integrator.cpp:41-59
//   System::BlockStep projects/nbody_core/include/ngrav/System.hpp:239-294 (rungs 249-265,
//                     substep loop 267-290, HalfKick 296-301)
//
// Adaptive dt is chosen on the device and stored in the device scalars, where the kick
// and drift kernels read it, so a step stays one graph with no host round trip:
//   dt = clamp(eta min_i sqrt(eps / |a_i|), dt0 / 4096, dt0).
// min_i sqrt(eps / |a_i|) equals sqrt(eps / max_i |a_i|) bit for bit (correctly rounded
// division and square root are monotonic), so one FP64 max-reduction of |a|^2 suffices.
//
// Block steps: per-particle rungs from |a| (FP64), then 2^rmax substeps; the host drives
// the substeps (their count depends on the rungs).  The closing set of a substep is
// compacted in order (sorted positions for the tree walk, original indices for the
// direct sum) by a count / scan / scatter, so the force pass sees the same target
// batches on every run.
#include <algorithm>
#include <cmath>

#include "nbg_common.hpp"
#include "nbg_internal.hpp"

namespace nbg {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = kReduceBlocks;

__device__ double block_max_d(double v, double* s_red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) s_red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) t = fmax(t, s_red[w]);
    __syncthreads();
    return t;
}

// |a|^2 in FP64 in the reference's order ((ax^2 + ay^2) + az^2), from the FP32 accelerations
__device__ __forceinline__ double amag2(const float* ax, const float* ay, const float* az, int i) {
    const double x = ax[i], y = ay[i], z = az[i];
    return __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

__global__ __launch_bounds__(kRedThreads) void k_amax_part(const float* __restrict__ ax, const float* __restrict__ ay,
                                                           const float* __restrict__ az, int n,
                                                           double* __restrict__ part) {
    __shared__ double s_red[kRedThreads / 64];
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, amag2(ax, ay, az, i));
    const double t = block_max_d(m, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// One block: the step's dt into sc->dt_step, accumulated into sc->t_accum.
__global__ __launch_bounds__(kRedThreads) void k_adaptive_dt(const double* __restrict__ part, int nparts, double eta,
                                                             double eps, double dt0, DevScalars* __restrict__ sc) {
    __shared__ double s_red[kRedThreads / 64];
    double m = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) m = fmax(m, part[k]);
    const double a2 = block_max_d(m, s_red);
    if (threadIdx.x == 0) {
        double dt = dt0;
        const double a = __dsqrt_rn(a2);
        if (eps > 0.0 && a > 0.0) {
            const double tau = __dsqrt_rn(eps / a);
            dt = eta * tau;
            // std::clamp(dt, dt0 / 4096, dt0)
            const double lo = dt0 / 4096.0;
            dt = (dt < lo) ? lo : ((dt0 < dt) ? dt0 : dt);
        }
        sc->dt_step = dt;
        sc->t_accum += dt;
    }
}

// Kick / kick+drift with the device-chosen dt (same arithmetic as the fixed-dt kernels:
// hdt = (float)(0.5 dt), dt rounded to float).
__global__ __launch_bounds__(256) void k_kick_drift_dev(float* __restrict__ x, float* __restrict__ y,
                                                        float* __restrict__ z, float* __restrict__ vx,
                                                        float* __restrict__ vy, float* __restrict__ vz,
                                                        const float* __restrict__ ax, const float* __restrict__ ay,
                                                        const float* __restrict__ az, int n,
                                                        const DevScalars* __restrict__ sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = sc->dt_step;
    const float hdt = static_cast<float>(0.5 * d), dt = static_cast<float>(d);
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

__global__ __launch_bounds__(256) void k_kick_dev(float* __restrict__ vx, float* __restrict__ vy,
                                                  float* __restrict__ vz, const float* __restrict__ ax,
                                                  const float* __restrict__ ay, const float* __restrict__ az, int n,
                                                  const DevScalars* __restrict__ sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float hdt = static_cast<float>(0.5 * sc->dt_step);
    vx[i] = vx[i] + hdt * ax[i];
    vy[i] = vy[i] + hdt * ay[i];
    vz[i] = vz[i] + hdt * az[i];
}

void launch_adaptive_dt(nbg_ctx* c, const nbg_params& p) {
    const int nb = std::min(kRedBlocks, std::max(1, grid_for(c->n, kRedThreads)));
    hipLaunchKernelGGL(k_amax_part, dim3(nb), dim3(kRedThreads), 0, c->stream, c->ax, c->ay, c->az, c->n,
                       c->red_part);
    hipLaunchKernelGGL(k_adaptive_dt, dim3(1), dim3(kRedThreads), 0, c->stream, c->red_part, nb, p.eta,
                       std::sqrt(p.eps2), p.dt, c->dsc);
}

void launch_kick_drift_dev(nbg_ctx* c) {
    hipLaunchKernelGGL(k_kick_drift_dev, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->x, c->y, c->z, c->vx,
                       c->vy, c->vz, c->ax, c->ay, c->az, c->n, c->dsc);
}

void launch_kick_dev(nbg_ctx* c) {
    hipLaunchKernelGGL(k_kick_dev, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->vx, c->vy, c->vz, c->ax,
                       c->ay, c->az, c->n, c->dsc);
}

// ---------------------------------------------------------------------------- block steps
// Rung of each particle (System.hpp:249-265): tau = eta sqrt(eps / |a|); k = the smallest
// power-of-2 subdivision of the coarse step `base` with base / 2^k <= tau, i.e.
// ceil(log2(base / tau)), clamped to [0, max_rung]; 0 when |a| = 0 or eps = 0.
__global__ __launch_bounds__(kRedThreads) void k_rungs(const float* __restrict__ ax, const float* __restrict__ ay,
                                                       const float* __restrict__ az, int n, double eta, double eps,
                                                       double base, int max_rung, int32_t* __restrict__ rung,
                                                       double* __restrict__ part) {
    __shared__ double s_red[kRedThreads / 64];
    double mk = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double a = __dsqrt_rn(amag2(ax, ay, az, i));
        int k = 0;
        if (a > 0.0 && eps > 0.0) {
            const double tau = eta * __dsqrt_rn(eps / a);
            if (tau < base) k = min(max_rung, static_cast<int>(ceil(log2(base / tau))));
        }
        rung[i] = k;
        mk = fmax(mk, static_cast<double>(k));
    }
    const double t = block_max_d(mk, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kRedThreads) void k_rung_max(const double* __restrict__ part, int nparts,
                                                          DevScalars* __restrict__ sc) {
    __shared__ double s_red[kRedThreads / 64];
    double m = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) m = fmax(m, part[k]);
    const double t = block_max_d(m, s_red);
    if (threadIdx.x == 0) sc->rung_max = static_cast<int32_t>(t);
}

// Half kick of the particles whose rung step starts (opening) or ends (closing) at
// substep `sub`: h = 0.5 base / 2^rung (HalfKick, System.hpp:296-301).
__global__ __launch_bounds__(256) void k_block_kick(float* __restrict__ vx, float* __restrict__ vy,
                                                    float* __restrict__ vz, const float* __restrict__ ax,
                                                    const float* __restrict__ ay, const float* __restrict__ az,
                                                    const int32_t* __restrict__ rung, int n, int sub, int rmax,
                                                    double base, int closing) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = rung[i];
    const int units = 1 << (rmax - r);
    const bool go = closing ? ((sub + 1) % units == 0) : (sub % units == 0);
    if (!go) return;
    const float h = static_cast<float>(0.5 * (base / static_cast<double>(1 << r)));
    vx[i] = vx[i] + h * ax[i];
    vy[i] = vy[i] + h * ay[i];
    vz[i] = vz[i] + h * az[i];
}

__global__ __launch_bounds__(256) void k_drift(float* __restrict__ x, float* __restrict__ y, float* __restrict__ z,
                                               const float* __restrict__ vx, const float* __restrict__ vy,
                                               const float* __restrict__ vz, int n, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = x[i] + dt * vx[i];
    y[i] = y[i] + dt * vy[i];
    z[i] = z[i] + dt * vz[i];
}

// Closing set of substep `sub` (System.hpp:279-283), compacted in order.  With `perm`,
// entry p of the scan is sorted position p (particle perm[p]) and the list holds sorted
// positions; without, it holds original indices.
constexpr int kListTile = 1024;  // positions per block (256 threads x 4)

__device__ __forceinline__ bool closing_at(const int32_t* rung, int i, int sub, int rmax) {
    return (sub + 1) % (1 << (rmax - rung[i])) == 0;
}

__global__ __launch_bounds__(256) void k_active_count(const int32_t* __restrict__ rung,
                                                      const uint32_t* __restrict__ perm, int n, int sub, int rmax,
                                                      uint32_t* __restrict__ blk_cnt) {
    __shared__ uint32_t s_w[4];
    uint32_t cnt = 0;
    for (int k = 0; k < 4; ++k) {
        const int p = blockIdx.x * kListTile + k * 256 + threadIdx.x;
        const bool f = p < n && closing_at(rung, perm ? static_cast<int>(perm[p]) : p, sub, rmax);
        cnt += static_cast<uint32_t>(__popcll(ballot(f)));
    }
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) blk_cnt[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// One block: exclusive scan of the block counts in place; the total goes to sc->tcount
// and is added to sc->particle_evals.
__global__ __launch_bounds__(1024) void k_active_scan(uint32_t* __restrict__ blk_cnt, int nblk,
                                                      DevScalars* __restrict__ sc) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < nblk; base += 1024) {
        const int k = base + static_cast<int>(threadIdx.x);
        const uint32_t v = (k < nblk) ? blk_cnt[k] : 0u;
        const uint32_t incl = wave_incl_scan(v);
        if (lane_id() == 63) s_w[threadIdx.x >> 6] = incl;
        __syncthreads();
        uint32_t woff = 0;
        for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); ++w) woff += s_w[w];
        const uint32_t carry = s_carry;
        if (k < nblk) blk_cnt[k] = carry + woff + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = carry + woff + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sc->tcount = s_carry;
        sc->particle_evals += s_carry;
    }
}

__global__ __launch_bounds__(256) void k_active_scatter(const int32_t* __restrict__ rung,
                                                        const uint32_t* __restrict__ perm, int n, int sub, int rmax,
                                                        const uint32_t* __restrict__ blk_off,
                                                        int32_t* __restrict__ list) {
    __shared__ uint32_t s_w[4][4];
    bool f[4];
    uint64_t m[4];
    for (int k = 0; k < 4; ++k) {
        const int p = blockIdx.x * kListTile + k * 256 + threadIdx.x;
        f[k] = p < n && closing_at(rung, perm ? static_cast<int>(perm[p]) : p, sub, rmax);
        m[k] = ballot(f[k]);
        if (lane_id() == 0) s_w[k][threadIdx.x >> 6] = static_cast<uint32_t>(__popcll(m[k]));
    }
    __syncthreads();
    // order: chunk k (256 positions) before chunk k+1; inside a chunk, wave by wave
    uint32_t off = blk_off[blockIdx.x];
    for (int k = 0; k < 4; ++k) {
        uint32_t woff = 0;
        for (int w = 0; w < 4; ++w) woff += (w < static_cast<int>(threadIdx.x >> 6)) ? s_w[k][w] : 0u;
        const int p = blockIdx.x * kListTile + k * 256 + threadIdx.x;
        if (f[k]) list[off + woff + static_cast<uint32_t>(mbcnt(m[k]))] = p;
        off += s_w[k][0] + s_w[k][1] + s_w[k][2] + s_w[k][3];
    }
}

void launch_rungs(nbg_ctx* c, const nbg_params& p) {
    const int nb = std::min(kRedBlocks, std::max(1, grid_for(c->n, kRedThreads)));
    hipLaunchKernelGGL(k_rungs, dim3(nb), dim3(kRedThreads), 0, c->stream, c->ax, c->ay, c->az, c->n, p.eta,
                       std::sqrt(p.eps2), p.dt, p.block_max_rung, c->rung, c->red_part);
    hipLaunchKernelGGL(k_rung_max, dim3(1), dim3(kRedThreads), 0, c->stream, c->red_part, nb, c->dsc);
}

void launch_block_kick(nbg_ctx* c, const nbg_params& p, int sub, int rmax, bool closing) {
    hipLaunchKernelGGL(k_block_kick, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->vx, c->vy, c->vz, c->ax,
                       c->ay, c->az, c->rung, c->n, sub, rmax, p.dt, closing ? 1 : 0);
}

void launch_drift(nbg_ctx* c, float dt) {
    hipLaunchKernelGGL(k_drift, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->x, c->y, c->z, c->vx, c->vy,
                       c->vz, c->n, dt);
}

void launch_active_list(nbg_ctx* c, int sub, int rmax, bool sorted) {
    const int nblk = std::max(1, grid_for(c->n, kListTile));
    const uint32_t* perm = sorted ? c->perm : nullptr;
    hipLaunchKernelGGL(k_active_count, dim3(nblk), dim3(256), 0, c->stream, c->rung, perm, c->n, sub, rmax,
                       c->list_off);
    hipLaunchKernelGGL(k_active_scan, dim3(1), dim3(1024), 0, c->stream, c->list_off, nblk, c->dsc);
    hipLaunchKernelGGL(k_active_scatter, dim3(nblk), dim3(256), 0, c->stream, c->rung, perm, c->n, sub, rmax,
                       c->list_off, c->tlist);
}

}  // namespace nbg
