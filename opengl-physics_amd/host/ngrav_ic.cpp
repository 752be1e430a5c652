// ngrav_ic.cpp -- seeded equilibrium-sphere initial conditions (host C++), the inputs of
// every configuration of the gravity path.  Restates the reference samplers with the same
// engine (libstdc++ std::mt19937), distributions and draw order:
//   Plummer    projects/nbody_core/include/ngrav/ic/Plummer.hpp:26-112
//   Hernquist  projects/nbody_core/include/ngrav/ic/Hernquist.hpp:51-181
//   King(W0)   projects/nbody_core/include/ngrav/ic/King.hpp:58-211
//   Cluster / ColdUniform  projects/nbody_core/src/Scenarios.cpp:104-142, 201-214
// Where the reference writes `Vec<3> v(gg(rng), gg(rng), gg(rng))` (unspecified argument
// order) the GCC order -- right to left, z drawn first -- is used.  Built without FP
// contraction so the FP64 values are the same on every host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/ngrav_host.h"

namespace {

struct Out {
    double *x, *y, *z, *vx, *vy, *vz, *m;
};

// Subtract the mass-weighted mean position and velocity (sampling drift).
void Recentre(const Out& o, int n) {
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0}, tm = 0.0;
    for (int i = 0; i < n; ++i) {
        cp[0] += o.m[i] * o.x[i];
        cp[1] += o.m[i] * o.y[i];
        cp[2] += o.m[i] * o.z[i];
        cv[0] += o.m[i] * o.vx[i];
        cv[1] += o.m[i] * o.vy[i];
        cv[2] += o.m[i] * o.vz[i];
        tm += o.m[i];
    }
    for (int k = 0; k < 3; ++k) {
        cp[k] /= tm;
        cv[k] /= tm;
    }
    for (int i = 0; i < n; ++i) {
        o.x[i] -= cp[0];
        o.y[i] -= cp[1];
        o.z[i] -= cp[2];
        o.vx[i] -= cv[0];
        o.vy[i] -= cv[1];
        o.vz[i] -= cv[2];
    }
}

// Isotropic unit direction: cos(theta) uniform, phi uniform (three-draw order as the reference).
struct Dir {
    double x, y, z;
};
Dir RandomDir(std::mt19937& rng, std::uniform_real_distribution<double>& u) {
    const double cosT = 2.0 * u(rng) - 1.0;
    const double sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
    const double phi = 2.0 * M_PI * u(rng);
    return {sinT * std::cos(phi), sinT * std::sin(phi), cosT};
}

// ---------------------------------------------------------------- Hernquist Jeans table
double HernquistRho(double r) {
    const double rp1 = r + 1.0;
    return 1.0 / (2.0 * M_PI * r * rp1 * rp1 * rp1);
}
double HernquistMass(double r) {
    const double rp1 = r + 1.0;
    return r * r / (rp1 * rp1);
}

// sigma_r^2(r) / (GM/a) from the spherical Jeans equation on a log grid (trapezoid rule,
// integrated inward from r = 200 a), linear interpolation in log r.
class HernquistSigma {
public:
    static constexpr int kN = 2048;
    static constexpr double kRMin = 1e-4, kRMax = 200.0;
    HernquistSigma() {
        lmin_ = std::log(kRMin);
        lmax_ = std::log(kRMax);
        std::vector<double> r(kN + 1), f(kN + 1), tail(kN + 1, 0.0);
        for (int i = 0; i <= kN; ++i) r[i] = std::exp(lmin_ + (lmax_ - lmin_) * i / kN);
        for (int i = 0; i <= kN; ++i) f[i] = HernquistRho(r[i]) * HernquistMass(r[i]) / (r[i] * r[i]);
        for (int i = kN - 1; i >= 0; --i) tail[i] = tail[i + 1] + 0.5 * (f[i] + f[i + 1]) * (r[i + 1] - r[i]);
        s2_.resize(kN + 1);
        for (int i = 0; i <= kN; ++i) {
            const double rho = HernquistRho(r[i]);
            s2_[i] = (rho > 1e-300) ? tail[i] / rho : 0.0;
        }
    }
    double operator()(double r) const {
        r = std::max(kRMin, std::min(kRMax, r));
        const double t = (std::log(r) - lmin_) / (lmax_ - lmin_) * kN;
        const int i0 = std::min(kN - 1, static_cast<int>(t));
        const double fr = t - i0;
        return s2_[i0] * (1.0 - fr) + s2_[i0 + 1] * fr;
    }

private:
    double lmin_ = 0, lmax_ = 0;
    std::vector<double> s2_;
};

// ---------------------------------------------------------------- King profile
double KingG(double W) {
    if (W <= 0.0) return 0.0;
    const double sq = std::sqrt(W);
    return std::exp(W) * std::erf(sq) - std::sqrt(4.0 * W / M_PI) * (1.0 + 2.0 * W / 3.0);
}

// Dimensionless King ODE W'' + (2/x) W' = -9 g(W)/g(W0) by RK4 (h = 2e-3) from the series
// start W = W0 - 1.5 x^2, tabulating r, the cumulative mass shape and the Jeans sigma^2.
struct KingProfile {
    std::vector<double> r, cum, rho, s2;
    KingProfile(double w0, double r0, double G, double M) {
        const double g0 = KingG(w0);
        const double h = 2e-3, xe = 1e-4;
        std::vector<double> x{0.0, xe}, W{w0, w0 - 1.5 * xe * xe};
        auto deriv = [&](double xx, double ww, double wp, double& dW, double& dWp) {
            dW = wp;
            dWp = -9.0 * KingG(ww) / g0 - ((xx > 1e-12) ? (2.0 / xx) * wp : 0.0);
        };
        double xc = xe, Wc = W.back(), Wpc = -3.0 * xe;
        while (Wc > 0.0 && x.size() < 200000) {
            double k1, l1, k2, l2, k3, l3, k4, l4;
            deriv(xc, Wc, Wpc, k1, l1);
            deriv(xc + h / 2, Wc + h / 2 * k1, Wpc + h / 2 * l1, k2, l2);
            deriv(xc + h / 2, Wc + h / 2 * k2, Wpc + h / 2 * l2, k3, l3);
            deriv(xc + h, Wc + h * k3, Wpc + h * l3, k4, l4);
            Wc = Wc + (h / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4);
            Wpc = Wpc + (h / 6.0) * (l1 + 2 * l2 + 2 * l3 + l4);
            xc += h;
            x.push_back(xc);
            W.push_back(std::max(Wc, 0.0));
        }
        const size_t np = x.size();
        r.resize(np);
        rho.resize(np);
        for (size_t i = 0; i < np; ++i) {
            r[i] = x[i] * r0;
            rho[i] = KingG(W[i]) / g0;
        }
        cum.assign(np, 0.0);
        for (size_t i = 1; i < np; ++i) {
            const double f0 = rho[i - 1] * r[i - 1] * r[i - 1], f1 = rho[i] * r[i] * r[i];
            cum[i] = cum[i - 1] + 0.5 * (f0 + f1) * (r[i] - r[i - 1]);
        }
        const double tot = cum.back();
        std::vector<double> mp(np), tail(np, 0.0);
        for (size_t i = 0; i < np; ++i) mp[i] = M * cum[i] / tot;
        for (size_t i = np - 1; i-- > 0;) {
            const double a = r[i], b = r[i + 1];
            const double f0 = (a > 1e-300) ? rho[i] * G * mp[i] / (a * a) : 0.0;
            const double f1 = rho[i + 1] * G * mp[i + 1] / (b * b);
            tail[i] = tail[i + 1] + 0.5 * (f0 + f1) * (b - a);
        }
        s2.resize(np);
        for (size_t i = 0; i < np; ++i) s2[i] = (rho[i] > 1e-12) ? tail[i] / rho[i] : 0.0;
    }
    static double Lerp(const std::vector<double>& xs, const std::vector<double>& ys, double v) {
        const auto it = std::lower_bound(xs.begin(), xs.end(), v);
        size_t i1 = static_cast<size_t>(it - xs.begin());
        i1 = std::clamp(i1, size_t{1}, xs.size() - 1);
        const size_t i0 = i1 - 1;
        const double fr = (xs[i1] > xs[i0]) ? (v - xs[i0]) / (xs[i1] - xs[i0]) : 0.0;
        return ys[i0] * (1.0 - fr) + ys[i1] * fr;
    }
    double SampleRadius(double x1) const { return Lerp(cum, r, x1 * cum.back()); }
    double Sigma2(double rr) const { return Lerp(r, s2, rr); }
};

}  // namespace

extern "C" {

int ngh_ic_plummer(int n, uint32_t seed, double totalMass, double scaleRadius, double G, double* x, double* y,
                   double* z, double* vx, double* vy, double* vz, double* m) {
    if (n < 1 || !x || !y || !z || !vx || !vy || !vz || !m) return -1;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double mi = totalMass / n;
    const double vUnit = std::sqrt(G * totalMass / scaleRadius);
    for (int i = 0; i < n; ++i) {
        const double x1 = u(rng);
        const double r = std::min(1.0 / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0), 20.0);
        const Dir d = RandomDir(rng, u);
        const double rr = r * scaleRadius;
        x[i] = d.x * rr;
        y[i] = d.y * rr;
        z[i] = d.z * rr;
        double q, g;
        do {
            q = u(rng);
            g = u(rng) * 0.1;
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double speed = q * (std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25));
        const Dir v = RandomDir(rng, u);
        const double sv = speed * vUnit;
        vx[i] = v.x * sv;
        vy[i] = v.y * sv;
        vz[i] = v.z * sv;
        m[i] = mi;
    }
    Recentre({x, y, z, vx, vy, vz, m}, n);
    return 0;
}

int ngh_ic_hernquist(int n, uint32_t seed, double totalMass, double scaleRadius, double G, double* x, double* y,
                     double* z, double* vx, double* vy, double* vz, double* m) {
    if (n < 1 || !x || !y || !z || !vx || !vy || !vz || !m) return -1;
    static const HernquistSigma sigma;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double mi = totalMass / n;
    const double vUnit = std::sqrt(G * totalMass / scaleRadius);
    for (int i = 0; i < n; ++i) {
        const double x1 = u(rng);
        const double sq = std::sqrt(x1);
        const double r = std::min(sq / (1.0 - sq), 40.0);
        const Dir d = RandomDir(rng, u);
        x[i] = d.x * (r * scaleRadius);
        y[i] = d.y * (r * scaleRadius);
        z[i] = d.z * (r * scaleRadius);
        std::normal_distribution<double> gg(0.0, std::sqrt(std::max(0.0, sigma(r))));
        const double gz = gg(rng), gy = gg(rng), gx = gg(rng);
        vx[i] = gx * vUnit;
        vy[i] = gy * vUnit;
        vz[i] = gz * vUnit;
        m[i] = mi;
    }
    Recentre({x, y, z, vx, vy, vz, m}, n);
    return 0;
}

int ngh_ic_king(int n, uint32_t seed, double totalMass, double coreRadius, double w0, double G, double* x,
                double* y, double* z, double* vx, double* vy, double* vz, double* m) {
    if (n < 1 || !x || !y || !z || !vx || !vy || !vz || !m || !(w0 > 0.0)) return -1;
    const KingProfile prof(w0, coreRadius, G, totalMass);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double mi = totalMass / n;
    for (int i = 0; i < n; ++i) {
        const double r = prof.SampleRadius(u(rng));
        const Dir d = RandomDir(rng, u);
        x[i] = d.x * r;
        y[i] = d.y * r;
        z[i] = d.z * r;
        std::normal_distribution<double> gg(0.0, std::sqrt(std::max(0.0, prof.Sigma2(r))));
        const double gz = gg(rng), gy = gg(rng), gx = gg(rng);
        vx[i] = gx;
        vy[i] = gy;
        vz[i] = gz;
        m[i] = mi;
    }
    Recentre({x, y, z, vx, vy, vz, m}, n);
    return 0;
}

// Cluster sizes and kinds of a batched ensemble: N_k = round(exp(U[ln nmin, ln nmax])),
// U from std::uniform_real_distribution<double> over std::mt19937(seed), drawn k = 0, 1, ...;
// kind_k = k % 2 (0 Plummer, 1 King).  Returns the total particle count, or -1.
int ngh_ic_cluster(int n, uint32_t seed, int cold, double* x, double* y, double* z, double* vx, double* vy,
                   double* vz, double* m) {
    // Cluster<3> (Scenarios.cpp:104-142): uniform sphere R = 1, M = 1, G = 1, Gaussian
    // velocities of dispersion 0.3 sqrt(G M / R) from one normal_distribution for the whole
    // run; ColdUniform (Scenarios.cpp:201-214) draws the same and zeroes the velocities.
    if (n < 1 || !x || !y || !z || !vx || !vy || !vz || !m) return -1;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::normal_distribution<double> gauss(0.0, 1.0);
    const double G = 1.0, R = 1.0, M = 1.0;
    const double vScale = std::sqrt(G * M / R);
    for (int i = 0; i < n; ++i) {
        const double radius = R * std::cbrt(u(rng));
        const Dir d = RandomDir(rng, u);
        x[i] = d.x * radius;
        y[i] = d.y * radius;
        z[i] = d.z * radius;
        const double gx = 0.3 * vScale * gauss(rng);
        const double gy = 0.3 * vScale * gauss(rng);
        const double gz = 0.3 * vScale * gauss(rng);
        vx[i] = cold ? 0.0 : gx;
        vy[i] = cold ? 0.0 : gy;
        vz[i] = cold ? 0.0 : gz;
        m[i] = M / n;
    }
    return 0;
}

int64_t ngh_ensemble_sizes(int32_t nclusters, uint32_t seed, int64_t nmin, int64_t nmax, int64_t* sizes,
                           int32_t* kinds) {
    if (nclusters < 1 || nmin < 1 || nmax < nmin || !sizes) return -1;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(std::log(static_cast<double>(nmin)), std::log(static_cast<double>(nmax)));
    int64_t total = 0;
    for (int32_t k = 0; k < nclusters; ++k) {
        const int64_t nk = std::llround(std::exp(u(rng)));
        sizes[k] = std::min(nmax, std::max(nmin, nk));
        if (kinds) kinds[k] = k % 2;
        total += sizes[k];
    }
    return total;
}

}  // extern "C"
