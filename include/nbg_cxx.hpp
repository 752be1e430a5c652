// nbg_cxx.hpp -- header-only C++ layer over the C ABI of include/nbg.h.
//
// Gives a C++ caller of the reference's nbody_core the shapes it already uses:
//   * nbg::Context        RAII owner of an nbg_ctx (one device, one stream)
//   * nbg::Params         StepParams-like value (projects/nbody_core/include/ngrav/Solvers.hpp:22-40)
//   * nbg::AccelInto      fill out.ax/ay/az from any SoA-like `in` (fields x,y,z,m), G included --
//                         the contract of ngrav::AccelFn<D> (projects/nbody_core/include/ngrav/Integrator.hpp:13-14)
//   * nbg::MakeAccelFn    a callable with AccelFn's signature, to register through
//                         System<3>::SetAuxSolver (projects/nbody_core/include/ngrav/System.hpp:40)
//   * nbg::BarnesHut / nbg::Direct
//                         the argument lists of ngrav::gpu::ComputeAccelGpuBarnesHut /
//                         ComputeAccelGpuDirect (projects/nbody_core/include/ngrav/gpu/GpuBarnesHut.hpp:40-42,
//                         GpuDirect.hpp:32-33), writing FP64 accelerations into a SoA-like `out`.
// Errors are reported as std::runtime_error, the reference's style (Scenarios.cpp:296).
#pragma once

#include <cmath>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "nbg.h"

namespace nbg {

struct Params {
    int solver = NBG_SOLVER_BARNES_HUT;
    int multipole = NBG_QUADRUPOLE;  // System<3> always adds quadrupoles (System.hpp:59)
    double G = 1.0;
    double eps = 0.05;               // Softening::eps: Plummer length or spline h
    int softening = NBG_SOFT_PLUMMER;
    double theta = 0.5;
    int mac = NBG_MAC_GEOMETRIC;     // MacParams::kind
    double alpha = 0.001;            // MacParams::alpha
    double dt = 1e-3;
    bool adaptive = false;           // StepParams::adaptive / eta / block / blockMaxRung
    double eta = 0.03;
    bool block = false;
    int block_max_rung = 8;
    bool kahan = true;

    nbg_params c() const {
        nbg_params p{};
        p.solver = solver;
        p.multipole = multipole;
        p.G = G;
        p.eps2 = eps * eps;
        p.theta = theta;
        p.dt = dt;
        p.flags = kahan ? NBG_FLAG_KAHAN : 0u;
        p.softening = softening;
        p.mac = mac;
        p.alpha = alpha;
        p.adaptive = adaptive ? 1 : 0;
        p.eta = eta;
        p.block = block ? 1 : 0;
        p.block_max_rung = block_max_rung;
        return p;
    }
};

class Context {
public:
    explicit Context(long long max_n, int device = 0) : h_(nbg_create(device, max_n)) {
        if (!h_) throw std::runtime_error("nbg_create failed (device " + std::to_string(device) + ")");
    }
    ~Context() { nbg_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    nbg_ctx* get() const { return h_; }
    void check(int rc, const char* what) const {
        if (rc != NBG_OK) throw std::runtime_error(std::string(what) + ": " + nbg_last_error(h_));
    }

private:
    nbg_ctx* h_;
};

namespace detail {
template <class V>
std::vector<float> to_f32(const V& v, size_t n) {
    std::vector<float> f(n);
    for (size_t i = 0; i < n; ++i) f[i] = static_cast<float>(v[i]);
    return f;
}
// does T carry accelerations (SoA<3>: ax, ay, az) or only positions (PosMassView<3>)?
template <class T, class = void>
struct has_accel : std::false_type {};
template <class T>
struct has_accel<T, std::void_t<decltype(std::declval<const T&>().ax.size())>> : std::true_type {};
}  // namespace detail

// Upload the positions/masses of `in`, evaluate, and write FP64 accelerations into `out`
// (resized to n, as SoA::ResizeAccel does).  In/Out: anything with x, y, z, m (and
// ax, ay, az on `out`) indexable containers of doubles -- SoA<3> or PosMassView<3>.
template <class In, class Out>
nbg_stats AccelInto(Context& ctx, const In& in, const Params& prm, Out& out) {
    const size_t n = in.x.size();
    const auto x = detail::to_f32(in.x, n), y = detail::to_f32(in.y, n), z = detail::to_f32(in.z, n),
               m = detail::to_f32(in.m, n);
    ctx.check(nbg_upload(ctx.get(), static_cast<int64_t>(n), x.data(), y.data(), z.data(), m.data(), nullptr,
                         nullptr, nullptr),
              "nbg_upload");
    // the relative MAC reads the previous |a|: the accelerations `in` carries (a(t) when
    // called from LeapfrogStep, Integrator.cpp:28), zero when it has none
    if constexpr (detail::has_accel<In>::value) {
        if (prm.mac == NBG_MAC_RELATIVE && in.ax.size() == n) {
            const auto ax = detail::to_f32(in.ax, n), ay = detail::to_f32(in.ay, n), az = detail::to_f32(in.az, n);
            ctx.check(nbg_upload_accel(ctx.get(), ax.data(), ay.data(), az.data()), "nbg_upload_accel");
        }
    }
    const nbg_params p = prm.c();
    nbg_stats st{};
    ctx.check(nbg_accel(ctx.get(), &p, &st), "nbg_accel");
    std::vector<float> ax(n), ay(n), az(n);
    ctx.check(nbg_download(ctx.get(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ax.data(), ay.data(),
                           az.data()),
              "nbg_download");
    out.ax.assign(ax.begin(), ax.end());
    out.ay.assign(ay.begin(), ay.end());
    out.az.assign(az.begin(), az.end());
    return st;
}

// A callable with the signature of ngrav::AccelFn<3>: void(const SoA&, const StepParams&, SoA&).
// The StepParams fields read are G, soft (kind, eps) and mac (kind, theta, alpha); the
// solver is fixed at creation.  SofteningKind and MacKind number their values as
// NBG_SOFT_* and NBG_MAC_* do (Softening.hpp:30-33, Mac.hpp:13-16).
template <class SoA, class StepParams>
std::function<void(const SoA&, const StepParams&, SoA&)> MakeAccelFn(std::shared_ptr<Context> ctx,
                                                                      int solver = NBG_SOLVER_BARNES_HUT,
                                                                      int multipole = NBG_QUADRUPOLE) {
    return [ctx, solver, multipole](const SoA& in, const StepParams& sp, SoA& out) {
        Params p;
        p.solver = solver;
        p.multipole = multipole;
        p.G = sp.G;
        p.eps = sp.soft.eps;
        p.softening = static_cast<int>(sp.soft.kind);
        p.theta = sp.mac.theta;
        p.mac = static_cast<int>(sp.mac.kind);
        p.alpha = sp.mac.alpha;
        AccelInto(*ctx, in, p, out);
    };
}

// ComputeAccelGpuBarnesHut's argument list.  treeMaxDepth / treeMaxCellsFactor size the
// reference's GL pools; this tree has a fixed depth cap of 40 and a pool bounded for any
// input, so they are accepted and ignored.
template <class View, class Out, class Stats = nbg_stats>
void BarnesHut(Context& ctx, const View& pts, double G, double eps2, double theta, Out& out,
               Stats* stats = nullptr, int /*treeMaxDepth*/ = 16, double /*treeMaxCellsFactor*/ = 4.0) {
    Params p;
    p.G = G;
    p.eps = std::sqrt(eps2);
    p.theta = theta;
    const nbg_stats st = AccelInto(ctx, pts, p, out);
    if (stats) *stats = st;
}

template <class View, class Out>
void Direct(Context& ctx, const View& pts, double G, double eps2, Out& out) {
    Params p;
    p.solver = NBG_SOLVER_DIRECT;
    p.G = G;
    p.eps = std::sqrt(eps2);
    AccelInto(ctx, pts, p, out);
}

}  // namespace nbg
