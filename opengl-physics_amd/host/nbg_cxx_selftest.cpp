// nbg_cxx_selftest -- exercises the C++ drop-in layer (include/nbg_cxx.hpp) the way the
// reference's own self-tests exercise nbody_core (projects/nbody_core/src/SelfTest.cpp):
// SoA-shaped state, StepParams-shaped parameters, an AccelFn-shaped force callback driving
// the kick-drift-kick step of Integrator.cpp:9-39.  Prints "name value (tol) ok|WRONG" and
// returns the verdict as the exit code.
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "../../include/nbg_cxx.hpp"
#include "../../include/ngrav_host.h"

namespace {

// Field names of ngrav::SoA<3> / StepParams (State.hpp:16-87, Solvers.hpp:22-40).
struct SoA {
    std::vector<double> x, y, z, vx, vy, vz, m, ax, ay, az;
    size_t Count() const { return x.size(); }
};
enum class SofteningKind { Plummer, Spline };  // Softening.hpp:30-33
enum class MacKind { Geometric, Relative };       // Mac.hpp:13-16
struct Soft {
    SofteningKind kind = SofteningKind::Plummer;
    double eps = 0.05;
    double eps2 = 0.05 * 0.05;
};
struct Mac {
    MacKind kind = MacKind::Geometric;
    double theta = 0.5;
    double alpha = 0.001;
};
struct StepParams {
    double G = 1.0;
    Soft soft;
    double dt = 1e-3;
    Mac mac;
};

bool ok = true;
void Check(const char* name, double value, double tol) {
    const bool good = std::isfinite(value) && std::abs(value) <= tol;
    std::printf("  %-48s %12.4e  (tol %.1e)  %s\n", name, value, tol, good ? "ok" : "WRONG");
    if (!good) ok = false;
}

template <class Fn>
void Leapfrog(SoA& s, const StepParams& sp, const Fn& accel, SoA& scratch) {
    const double hdt = 0.5 * sp.dt;
    for (size_t i = 0; i < s.Count(); ++i) {
        s.vx[i] += hdt * s.ax[i];
        s.vy[i] += hdt * s.ay[i];
        s.vz[i] += hdt * s.az[i];
    }
    for (size_t i = 0; i < s.Count(); ++i) {
        s.x[i] += sp.dt * s.vx[i];
        s.y[i] += sp.dt * s.vy[i];
        s.z[i] += sp.dt * s.vz[i];
    }
    accel(s, sp, scratch);
    s.ax = scratch.ax;
    s.ay = scratch.ay;
    s.az = scratch.az;
    for (size_t i = 0; i < s.Count(); ++i) {
        s.vx[i] += hdt * s.ax[i];
        s.vy[i] += hdt * s.ay[i];
        s.vz[i] += hdt * s.az[i];
    }
}

}  // namespace

int main() {
    auto ctx = std::make_shared<nbg::Context>(1 << 16);

    // [kepler, AccelFn] two-body e = 0.36 orbit over 2 periods (SelfTest.cpp:83-117 set-up)
    {
        std::printf("[kepler via AccelFn]\n");
        const double G = 1.0, m = 1.0, sep = 1.0, vRel = 0.8 * std::sqrt(2.0);
        const double period = 2.0 * M_PI * std::sqrt(sep * sep * sep / (G * 2.0 * m));
        SoA s;
        s.x = {-0.5, 0.5};
        s.y = {0, 0};
        s.z = {0, 0};
        s.vx = {0, 0};
        s.vy = {-vRel / 2, vRel / 2};
        s.vz = {0, 0};
        s.m = {m, m};
        StepParams sp;
        sp.soft.eps = 1e-6;
        sp.dt = period / 2000.0;
        auto accel = nbg::MakeAccelFn<SoA, StepParams>(ctx, NBG_SOLVER_DIRECT);
        SoA scratch;
        accel(s, sp, s);  // System::Prime
        double rMin = 1e30, rMax = 0;
        for (int k = 0; k < 2 * 2000; ++k) {
            Leapfrog(s, sp, accel, scratch);
            const double r = std::hypot(s.x[1] - s.x[0], s.y[1] - s.y[0]);
            rMin = std::min(rMin, r);
            rMax = std::max(rMax, r);
        }
        const double a = 1.0 / 1.36;
        // FP32 state on the device: position rounding bounds the orbit's accuracy
        Check("periapsis rel err", (rMin - a * 0.64) / (a * 0.64), 2e-4);
        Check("apoapsis rel err", rMax - 1.0, 2e-4);
    }

    // [ComputeAccelGpuBarnesHut shape] theta = 0 equals direct (SelfTest.cpp:43-74)
    {
        std::printf("[barnes-hut vs direct through the C++ layer]\n");
        const int n = 4096;
        SoA s;
        for (auto* v : {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.m}) v->resize(n);
        ngh_ic_plummer(n, 7, 1.0, 1.0, 1.0, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(),
                       s.vz.data(), s.m.data());
        SoA aD, aB;
        nbg::Direct(*ctx, s, 1.0, 1e-4, aD);
        nbg::BarnesHut(*ctx, s, 1.0, 1e-4, 0.0, aB);
        double worst = 0;
        for (int i = 0; i < n; ++i) {
            const double dx = aB.ax[i] - aD.ax[i], dy = aB.ay[i] - aD.ay[i], dz = aB.az[i] - aD.az[i];
            worst = std::max(worst, std::sqrt(dx * dx + dy * dy + dz * dz) /
                                        std::sqrt(aD.ax[i] * aD.ax[i] + aD.ay[i] * aD.ay[i] + aD.az[i] * aD.az[i]));
        }
        Check("max |da|/|a| theta=0 (FP32 sums)", worst, 1e-5);
        nbg_stats st{};
        nbg::BarnesHut(*ctx, s, 1.0, 1e-4, 0.5, aB, &st);
        double num = 0, den = 0;
        for (int i = 0; i < n; ++i) {
            num += std::pow(aB.ax[i] - aD.ax[i], 2) + std::pow(aB.ay[i] - aD.ay[i], 2) + std::pow(aB.az[i] - aD.az[i], 2);
            den += aD.ax[i] * aD.ax[i] + aD.ay[i] * aD.ay[i] + aD.az[i] * aD.az[i];
        }
        Check("rel RMS BH(theta=0.5, quad) vs direct", std::sqrt(num / den), 2e-3);
        Check("tree has nodes", st.n_nodes > 0 ? 0.0 : 1.0, 0.0);
    }
    std::printf("%s\n", ok ? "ALL OK" : "FAILURES");
    return ok ? 0 : 1;
}
