// nbody_oracle.cpp -- FP64 CPU restatement of the reference nbody_core gravity path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is compared
// against.  Only the test-suite, the smoke check and the benchmark's CPU-baseline
// leg may load it, and only as the checker or the timed CPU baseline, never as the
// computation that is shipped.  The product library (libnbg_hip.so) neither links
// nor loads it.
//
// Every function restates a reference function term for term (same operation
// order, FP64, no fused multiply-add when built with -ffp-contract=off).  Citations
// are relative to the reference repository root.
//
//   orc_ic_plummer          projects/nbody_core/include/ngrav/ic/Plummer.hpp:26-112
//   orc_ic_hernquist        projects/nbody_core/include/ngrav/ic/Hernquist.hpp:51-181
//   orc_ic_king             projects/nbody_core/include/ngrav/ic/King.hpp:58-211
//   orc_king_tidal_radius   projects/nbody_core/include/ngrav/ic/King.hpp:78-106
//   orc_tree_build          projects/nbody_core/src/Tree.cpp:33-203 (serial insertion)
//   quadrupoles             projects/nbody_core/src/Tree.cpp:205-254
//   orc_accel_bh            projects/nbody_core/src/Solvers.cpp:248-355 (grouped walk),
//                           pair kernel Solvers.cpp:19-38, quadrupole Solvers.cpp:147-179
//   orc_accel_bh_pp         projects/nbody_core/src/Solvers.cpp:181-227, 359-382
//   orc_accel_direct        projects/nbody_core/src/Solvers.cpp:77-106
//   orc_accel_bh_targets    projects/nbody_core/src/Solvers.cpp:384-405
//   softening kinds         Solvers.cpp:19-73, Softening.hpp:39-56, 120-139
//   orc_steps / orc_leapfrog projects/nbody_core/src/Integrator.cpp:9-39 with System::Step
//                           (System.hpp:89-108) and CaptureAOld (302-310)
//   orc_adaptive_dt         projects/nbody_core/src/Integrator.cpp:41-59
//   orc_rungs, orc_block_steps  System::BlockStep, System.hpp:239-301
//   orc_total_energy        projects/nbody_core/include/ngrav/System.hpp:119-164
//   orc_com / orc_angmom    projects/nbody_core/include/ngrav/System.hpp:166-193
//   orc_morton_keys         derived contract (not a reference function): the octant
//                           digits the insertion build of Tree.cpp:11-20, 96-99,
//                           114-145 assigns to a particle, level by level, from the
//                           same FP64 box-centre recurrence.
//
// Pinning (what checks each function beyond retyping it) is recorded in the
// tests: tests/test_oracle.py.  No function here is "parity unpinned" except the
// bit patterns of the IC samplers against a contraction-enabled reference build,
// which no fixture of the reference covers.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

namespace {

struct Tree {
    int ncrit = 8, maxDepth = 40;
    // node SoA in insertion order (Tree.hpp:44-58)
    std::vector<double> cx, cy, cz, half, comx, comy, comz, mass, maxR;
    std::vector<std::array<int, 8>> child;
    std::vector<int> childCount, parent, first, count, depth;
    std::vector<int> order;
    std::vector<double> Qxx, Qyy, Qxy, Qxz, Qyz;
    bool hasQ = false;
    std::vector<std::vector<int>> bucket;
    // canonical numbering: DFS pre-order, children in ascending octant
    std::vector<int> canon;    // insertion id -> canonical id
    std::vector<int> fromCanon; // canonical id -> insertion id
};

struct Pts {
    const double *x, *y, *z, *m;
    int n;
};

int OctantOf(double cx, double cy, double cz, double px, double py, double pz) {
    int oct = 0;
    if (px >= cx) oct |= 1;
    if (py >= cy) oct |= 2;
    if (pz >= cz) oct |= 4;
    return oct;
}

int NewNode(Tree& t, double cx, double cy, double cz, double h, int parent, int depth) {
    const int id = static_cast<int>(t.mass.size());
    t.cx.push_back(cx);
    t.cy.push_back(cy);
    t.cz.push_back(cz);
    t.half.push_back(h);
    t.comx.push_back(0.0);
    t.comy.push_back(0.0);
    t.comz.push_back(0.0);
    t.mass.push_back(0.0);
    std::array<int, 8> none;
    none.fill(-1);
    t.child.push_back(none);
    t.childCount.push_back(0);
    t.parent.push_back(parent);
    t.first.push_back(-1);
    t.count.push_back(0);
    t.depth.push_back(depth);
    t.bucket.emplace_back();
    return id;
}

void RootBox(const Pts& p, double c[3], double* halfOut) {
    double lo[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double hi[3] = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
                    std::numeric_limits<double>::lowest()};
    for (int i = 0; i < p.n; ++i) {
        const double v[3] = {p.x[i], p.y[i], p.z[i]};
        for (int k = 0; k < 3; ++k) {
            lo[k] = (v[k] < lo[k]) ? v[k] : lo[k];
            hi[k] = (hi[k] < v[k]) ? v[k] : hi[k];
        }
    }
    if (p.n == 0) {
        for (int k = 0; k < 3; ++k) lo[k] = hi[k] = 0.0;
    }
    double half = 0.5 * (hi[0] - lo[0]);
    half = std::max(half, 0.5 * (hi[1] - lo[1]));
    half = std::max(half, 0.5 * (hi[2] - lo[2]));
    if (!(half > 1e-12)) half = 1.0;
    for (int k = 0; k < 3; ++k) c[k] = 0.5 * (lo[k] + hi[k]);
    *halfOut = half * 1.001;
}

void Insert(Tree& t, const Pts& p, int node, int i, int depth);

void InsertIntoChild(Tree& t, const Pts& p, int node, int i, int depth) {
    const int oct = OctantOf(t.cx[node], t.cy[node], t.cz[node], p.x[i], p.y[i], p.z[i]);
    int ch = t.child[node][oct];
    if (ch == -1) {
        const double h = 0.5 * t.half[node];
        const double dx = (oct & 1) ? h : -h, dy = (oct & 2) ? h : -h, dz = (oct & 4) ? h : -h;
        ch = NewNode(t, t.cx[node] + dx, t.cy[node] + dy, t.cz[node] + dz, h, node, depth + 1);
        t.child[node][oct] = ch;
        t.childCount[node] += 1;
    }
    Insert(t, p, ch, i, depth + 1);
}

void Insert(Tree& t, const Pts& p, int node, int i, int depth) {
    const double mp = p.m[i];
    const double newMass = t.mass[node] + mp;
    t.comx[node] = (t.comx[node] * t.mass[node] + p.x[i] * mp) / newMass;
    t.comy[node] = (t.comy[node] * t.mass[node] + p.y[i] * mp) / newMass;
    t.comz[node] = (t.comz[node] * t.mass[node] + p.z[i] * mp) / newMass;
    t.mass[node] = newMass;
    if (t.childCount[node] == 0) {
        if (static_cast<int>(t.bucket[node].size()) < t.ncrit || depth >= t.maxDepth) {
            t.bucket[node].push_back(i);
            return;
        }
        std::vector<int> held;
        held.swap(t.bucket[node]);
        for (int q : held) InsertIntoChild(t, p, node, q, depth);
    }
    InsertIntoChild(t, p, node, i, depth);
}

void Finalize(Tree& t, const Pts& p) {
    const int nn = static_cast<int>(t.mass.size());
    t.order.clear();
    t.order.reserve(p.n);
    t.maxR.assign(nn, 0.0);
    t.canon.assign(nn, -1);
    t.fromCanon.clear();
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int node = stack.back();
        stack.pop_back();
        t.canon[node] = static_cast<int>(t.fromCanon.size());
        t.fromCanon.push_back(node);
        if (t.childCount[node] == 0) {
            t.first[node] = static_cast<int>(t.order.size());
            for (int q : t.bucket[node]) t.order.push_back(q);
            t.count[node] = static_cast<int>(t.bucket[node].size());
            continue;
        }
        for (int c = 7; c >= 0; --c)
            if (t.child[node][c] != -1) stack.push_back(t.child[node][c]);
    }
    t.bucket.clear();
    for (int nd = nn - 1; nd >= 0; --nd) {
        if (t.mass[nd] <= 0.0) continue;
        double r = 0.0;
        if (t.childCount[nd] == 0) {
            for (int k = t.first[nd]; k < t.first[nd] + t.count[nd]; ++k) {
                const int q = t.order[k];
                const double dx = p.x[q] - t.comx[nd], dy = p.y[q] - t.comy[nd], dz = p.z[q] - t.comz[nd];
                r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        } else {
            for (int c = 0; c < 8; ++c) {
                const int ch = t.child[nd][c];
                if (ch == -1) continue;
                const double dx = t.comx[ch] - t.comx[nd], dy = t.comy[ch] - t.comy[nd], dz = t.comz[ch] - t.comz[nd];
                r = std::max(r, t.maxR[ch] + std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        t.maxR[nd] = r;
    }
}

void Quadrupoles(Tree& t, const Pts& p) {
    const int nn = static_cast<int>(t.mass.size());
    t.Qxx.assign(nn, 0.0);
    t.Qyy.assign(nn, 0.0);
    t.Qxy.assign(nn, 0.0);
    t.Qxz.assign(nn, 0.0);
    t.Qyz.assign(nn, 0.0);
    t.hasQ = true;
    for (int nd = nn - 1; nd >= 0; --nd) {
        if (t.mass[nd] <= 0.0) continue;
        if (t.childCount[nd] == 0) {
            if (t.count[nd] <= 1) continue;
            for (int k = t.first[nd]; k < t.first[nd] + t.count[nd]; ++k) {
                const int q = t.order[k];
                const double dx = p.x[q] - t.comx[nd], dy = p.y[q] - t.comy[nd], dz = p.z[q] - t.comz[nd];
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double mp = p.m[q];
                t.Qxx[nd] += mp * (3.0 * dx * dx - d2);
                t.Qyy[nd] += mp * (3.0 * dy * dy - d2);
                t.Qxy[nd] += mp * 3.0 * dx * dy;
                t.Qxz[nd] += mp * 3.0 * dx * dz;
                t.Qyz[nd] += mp * 3.0 * dy * dz;
            }
            continue;
        }
        for (int c = 0; c < 8; ++c) {
            const int ch = t.child[nd][c];
            if (ch == -1 || t.mass[ch] <= 0.0) continue;
            const double dx = t.comx[ch] - t.comx[nd], dy = t.comy[ch] - t.comy[nd], dz = t.comz[ch] - t.comz[nd];
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double mc = t.mass[ch];
            t.Qxx[nd] += t.Qxx[ch] + mc * (3.0 * dx * dx - d2);
            t.Qyy[nd] += t.Qyy[ch] + mc * (3.0 * dy * dy - d2);
            t.Qxy[nd] += t.Qxy[ch] + mc * 3.0 * dx * dy;
            t.Qxz[nd] += t.Qxz[ch] + mc * 3.0 * dx * dz;
            t.Qyz[nd] += t.Qyz[ch] + mc * 3.0 * dy * dz;
        }
    }
}

// Softening::kind / eps / eps2 (Softening.hpp:120-139): spline stores eps = h, eps2 = h^2.
struct Soft {
    int spline;
    double eps;
    double eps2;
};
Soft MakeSoft(double eps2, int spline) { return Soft{spline, std::sqrt(eps2), eps2}; }

// SplineChi3D (Softening.hpp:39-45).
inline double SplineChi3D(double q) {
    if (q >= 2.0) return 1.0;
    if (q < 1.0) return q * q * q * (4.0 / 3.0 - 1.2 * q * q + 0.5 * q * q * q);
    const double t = 2.0 - q;
    const double t2 = t * t, t4 = t2 * t2;
    return 1.0 - 4.0 * (t4 / 4.0 - t4 * t / 5.0 + t4 * t2 / 24.0);
}

// SplinePotential3D (Softening.hpp:49-56).
inline double SplinePotential3D(double q) {
    if (q >= 2.0) return -1.0 / q;
    if (q < 1.0) return -(1.4 - (2.0 / 3.0) * q * q + 0.3 * q * q * q * q - 0.1 * q * q * q * q * q);
    const double t = 2.0 - q;
    const double t2 = t * t, t4 = t2 * t2;
    const double F = t4 / 4.0 - t4 * t / 5.0 + t4 * t2 / 24.0;
    return -((1.0 - 4.0 * F) / q + t4 / 2.0 - t4 * t / 5.0);
}

// Plummer pair kernel (Solvers.cpp:19-38); gmj = G*m_j.
inline void AccumPairPlummer(double dx, double dy, double dz, double gmj, double eps2, double& ax, double& ay,
                             double& az) {
    const double dist2 = dx * dx + dy * dy + dz * dz + eps2;
    const double inv = 1.0 / std::sqrt(dist2);
    const double inv3 = inv * inv * inv;
    const double s = gmj * inv3;
    ax += s * dx;
    ay += s * dy;
    az += s * dz;
}

// Spline pair kernel (Solvers.cpp:43-63), h = Softening::eps.
inline void AccumPairSpline(double dx, double dy, double dz, double gmj, double h, double& ax, double& ay,
                            double& az) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = std::sqrt(r2);
    if (r <= 1e-300) return;
    const double q = r / h;
    const double chi = SplineChi3D(q);
    const double s = gmj * chi / (r2 * r);
    ax += s * dx;
    ay += s * dy;
    az += s * dz;
}

// AccumPair (Solvers.cpp:65-73).
inline void AccumPair(double dx, double dy, double dz, double gmj, const Soft& soft, double& ax, double& ay,
                      double& az) {
    if (!soft.spline)
        AccumPairPlummer(dx, dy, dz, gmj, soft.eps2, ax, ay, az);
    else
        AccumPairSpline(dx, dy, dz, gmj, soft.eps, ax, ay, az);
}

// Traceless quadrupole term, unsoftened (Solvers.cpp:147-179).
inline void AddQuad(const Tree& t, int nd, double dx, double dy, double dz, double dist2, double G, double& ax,
                    double& ay, double& az) {
    if (!t.hasQ) return;
    const double Qxx = t.Qxx[nd], Qyy = t.Qyy[nd], Qzz = -(Qxx + Qyy);
    const double Qxy = t.Qxy[nd], Qxz = t.Qxz[nd], Qyz = t.Qyz[nd];
    const double Qdx = Qxx * dx + Qxy * dy + Qxz * dz;
    const double Qdy = Qxy * dx + Qyy * dy + Qyz * dz;
    const double Qdz = Qxz * dx + Qyz * dy + Qzz * dz;
    const double Qd2 = dx * Qdx + dy * Qdy + dz * Qdz;
    const double invR2 = 1.0 / dist2;
    const double invR = std::sqrt(invR2);
    const double invR5 = invR2 * invR2 * invR;
    const double invR7 = invR5 * invR2;
    const double c1 = -G * invR5;
    const double c2 = 2.5 * G * Qd2 * invR7;
    ax += c1 * Qdx + c2 * dx;
    ay += c1 * Qdy + c2 * dy;
    az += c1 * Qdz + c2 * dz;
}

// CollectGroupInteractions (Solvers.cpp:248-277).
void Collect(const Tree& t, int node, int group, double gx, double gy, double gz, double R, double theta2,
             std::vector<int>& far, std::vector<int>& near) {
    if (node == group) return;
    if (t.mass[node] <= 0.0) return;
    if (t.childCount[node] == 0) {
        near.push_back(node);
        return;
    }
    const double dx = t.comx[node] - gx, dy = t.comy[node] - gy, dz = t.comz[node] - gz;
    const double dist2 = dx * dx + dy * dy + dz * dz;
    const double size = 2.0 * t.half[node];
    const double worst = std::sqrt(dist2) - R;
    if (worst > 0.0 && size * size < theta2 * worst * worst) {
        far.push_back(node);
        return;
    }
    for (int c = 0; c < 8; ++c)
        if (t.child[node][c] != -1) Collect(t, t.child[node][c], group, gx, gy, gz, R, theta2, far, near);
}

// Per-particle walk, geometric or relative MAC (Solvers.cpp:181-227).
void WalkPP(const Tree& t, const Pts& p, int node, int i, double xi, double yi, double zi, double G, double theta2,
            const Soft& eps2, int relative, double alpha, double aOld, double& ax, double& ay, double& az,
            int64_t* cnt = nullptr) {
    if (t.mass[node] <= 0.0) return;
    if (t.childCount[node] == 0) {
        for (int k = t.first[node]; k < t.first[node] + t.count[node]; ++k) {
            const int q = t.order[k];
            if (q == i) continue;
            AccumPair(p.x[q] - xi, p.y[q] - yi, p.z[q] - zi, G * p.m[q], eps2, ax, ay, az);
            if (cnt) ++cnt[1];
        }
        return;
    }
    const double dx = t.comx[node] - xi, dy = t.comy[node] - yi, dz = t.comz[node] - zi;
    const double dist2 = dx * dx + dy * dy + dz * dz;
    const double size = 2.0 * t.half[node];
    bool accept;
    if (relative && aOld > 0.0)
        accept = (t.mass[node] * size * size) < (alpha * aOld * dist2 * dist2 / G);
    else
        accept = (size * size) < (theta2 * dist2);
    if (accept) {
        AccumPair(dx, dy, dz, G * t.mass[node], eps2, ax, ay, az);
        AddQuad(t, node, dx, dy, dz, dist2, G, ax, ay, az);
        if (cnt) ++cnt[0];
        return;
    }
    for (int c = 0; c < 8; ++c)
        if (t.child[node][c] != -1)
            WalkPP(t, p, t.child[node][c], i, xi, yi, zi, G, theta2, eps2, relative, alpha, aOld, ax, ay, az, cnt);
}

} // namespace

extern "C" {

// ---------------------------------------------------------------- initial conditions
// ic::Plummer<3> (Plummer.hpp:26-112).  Uses libstdc++ std::mt19937 and
// uniform_real_distribution<double>, exactly the reference's engine and distribution.
void orc_ic_plummer(int n, uint32_t seed, double totalMass, double scaleRadius, double G, double* x, double* y,
                    double* z, double* vx, double* vy, double* vz, double* m) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double mi = totalMass / n;
    const double vUnit = std::sqrt(G * totalMass / scaleRadius);
    for (int i = 0; i < n; ++i) {
        const double x1 = u(rng);
        double r = 1.0 / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0);
        r = std::min(r, 20.0);
        double cosT = 2.0 * u(rng) - 1.0;
        double sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
        double phi = 2.0 * M_PI * u(rng);
        const double rr = r * scaleRadius;
        x[i] = sinT * std::cos(phi) * rr;
        y[i] = sinT * std::sin(phi) * rr;
        z[i] = cosT * rr;
        double q, g;
        do {
            q = u(rng);
            g = u(rng) * 0.1;
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double vEsc = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
        const double speed = q * vEsc;
        cosT = 2.0 * u(rng) - 1.0;
        sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
        phi = 2.0 * M_PI * u(rng);
        const double sv = speed * vUnit;
        vx[i] = sinT * std::cos(phi) * sv;
        vy[i] = sinT * std::sin(phi) * sv;
        vz[i] = cosT * sv;
        m[i] = mi;
    }
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0}, tm = 0.0;
    for (int i = 0; i < n; ++i) {
        cp[0] += m[i] * x[i];
        cp[1] += m[i] * y[i];
        cp[2] += m[i] * z[i];
        cv[0] += m[i] * vx[i];
        cv[1] += m[i] * vy[i];
        cv[2] += m[i] * vz[i];
        tm += m[i];
    }
    for (int k = 0; k < 3; ++k) {
        cp[k] /= tm;
        cv[k] /= tm;
    }
    for (int i = 0; i < n; ++i) {
        x[i] -= cp[0];
        y[i] -= cp[1];
        z[i] -= cp[2];
        vx[i] -= cv[0];
        vy[i] -= cv[1];
        vz[i] -= cv[2];
    }
}

// ic::Hernquist<3> (Hernquist.hpp:131-181) with its Jeans table (Hernquist.hpp:51-127).
// The table is rebuilt per call (the reference keeps one static instance; the values are
// the same).  Velocity components: the reference writes Vec<3> v(gg(rng), gg(rng), gg(rng))
// (Hernquist.hpp:160), whose argument order C++ leaves unspecified; GCC on x86-64 evaluates
// it right to left, so the first normal draw is v.z, then v.y, then v.x.
void orc_ic_hernquist(int n, uint32_t seed, double totalMass, double scaleRadius, double G, double* x, double* y,
                      double* z, double* vx, double* vy, double* vz, double* m) {
    // HernquistSigmaTable (Hernquist.hpp:69-122): kN = 2048 log-spaced radii on [1e-4, 200]
    const int kN = 2048;
    const double kRMin = 1e-4, kRMax = 200.0;
    const double logMin = std::log(kRMin), logMax = std::log(kRMax);
    auto rhoUnit = [](double r) {  // Hernquist.hpp:54-57
        const double rp1 = r + 1.0;
        return 1.0 / (2.0 * M_PI * r * rp1 * rp1 * rp1);
    };
    auto massUnit = [](double r) {  // Hernquist.hpp:58-61
        const double rp1 = r + 1.0;
        return r * r / (rp1 * rp1);
    };
    std::vector<double> grid(kN + 1), integrand(kN + 1), tailInt(kN + 1, 0.0), sigma2(kN + 1);
    for (int i = 0; i <= kN; ++i) grid[i] = std::exp(logMin + (logMax - logMin) * i / kN);      // :80-83
    for (int i = 0; i <= kN; ++i)
        integrand[i] = rhoUnit(grid[i]) * massUnit(grid[i]) / (grid[i] * grid[i]);           // :84-91
    for (int i = kN - 1; i >= 0; --i)                                                           // :96-102
        tailInt[i] = tailInt[i + 1] + 0.5 * (integrand[i] + integrand[i + 1]) * (grid[i + 1] - grid[i]);
    for (int i = 0; i <= kN; ++i) {                                                             // :103-108
        const double rho = rhoUnit(grid[i]);
        sigma2[i] = (rho > 1e-300) ? (tailInt[i] / rho) : 0.0;
    }
    auto sigma2At = [&](double r) {  // Sigma2 (Hernquist.hpp:111-117)
        r = std::max(kRMin, std::min(kRMax, r));
        const double t = (std::log(r) - logMin) / (logMax - logMin) * kN;
        const int i0 = std::min(kN - 1, static_cast<int>(t));
        const double frac = t - i0;
        return sigma2[i0] * (1.0 - frac) + sigma2[i0 + 1] * frac;
    };

    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double mi = totalMass / n;
    const double vUnit = std::sqrt(G * totalMass / scaleRadius);
    for (int i = 0; i < n; ++i) {
        const double x1 = u(rng);                                       // :148-150
        const double sq = std::sqrt(x1);
        const double r = std::min(sq / (1.0 - sq), 40.0);
        const double cosT = 2.0 * u(rng) - 1.0;                         // :152-156
        const double sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
        const double phi = 2.0 * M_PI * u(rng);
        const double rs = r * scaleRadius;
        x[i] = sinT * std::cos(phi) * rs;
        y[i] = sinT * std::sin(phi) * rs;
        z[i] = cosT * rs;
        const double sigma = std::sqrt(std::max(0.0, sigma2At(r)));     // :158-161
        std::normal_distribution<double> gg(0.0, sigma);
        const double drawZ = gg(rng);
        const double drawY = gg(rng);
        const double drawX = gg(rng);
        vx[i] = drawX * vUnit;
        vy[i] = drawY * vUnit;
        vz[i] = drawZ * vUnit;
        m[i] = mi;
    }
    // recentre (Hernquist.hpp:165-179)
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0}, tm = 0.0;
    for (int i = 0; i < n; ++i) {
        cp[0] += m[i] * x[i];
        cp[1] += m[i] * y[i];
        cp[2] += m[i] * z[i];
        cv[0] += m[i] * vx[i];
        cv[1] += m[i] * vy[i];
        cv[2] += m[i] * vz[i];
        tm += m[i];
    }
    for (int k = 0; k < 3; ++k) {
        cp[k] /= tm;
        cv[k] /= tm;
    }
    for (int i = 0; i < n; ++i) {
        x[i] -= cp[0];
        y[i] -= cp[1];
        z[i] -= cp[2];
        vx[i] -= cv[0];
        vy[i] -= cv[1];
        vz[i] -= cv[2];
    }
}

namespace {
// KingG (King.hpp:60-64): g(W) = e^W erf(sqrt W) - sqrt(4W/pi) (1 + 2W/3).
double OrcKingG(double W) {
    if (W <= 0.0) return 0.0;
    const double sq = std::sqrt(W);
    return std::exp(W) * std::erf(sq) - std::sqrt(4.0 * W / M_PI) * (1.0 + 2.0 * W / 3.0);
}

// KingProfile (King.hpp:71-164): the RK4 solution of W'' + (2/x) W' = -9 g(W)/g(W0) and the
// tables derived from it.
struct OrcKingProfile {
    double xt = 0.0;
    std::vector<double> r, cumMassShape, rhoShape, sigma2;
    OrcKingProfile(double w0, double r0, double G, double totalMass) {
        const double g0 = OrcKingG(w0);
        const double h = 2e-3, xEps = 1e-4;                                       // :80-81
        std::vector<double> xs{0.0, xEps};                                        // :83-85
        std::vector<double> Ws{w0, w0 - 1.5 * xEps * xEps};
        double xc = xEps, Wc = Ws.back(), Wpc = -3.0 * xEps;
        auto dWp = [&](double xx, double WW, double WWp) {                       // :87-90
            return -9.0 * OrcKingG(WW) / g0 - ((xx > 1e-12) ? (2.0 / xx) * WWp : 0.0);
        };
        while (Wc > 0.0 && xs.size() < 200000) {                                  // :93-105
            const double k1W = Wpc, k1P = dWp(xc, Wc, Wpc);
            const double k2W = Wpc + h / 2 * k1P, k2P = dWp(xc + h / 2, Wc + h / 2 * k1W, Wpc + h / 2 * k1P);
            const double k3W = Wpc + h / 2 * k2P, k3P = dWp(xc + h / 2, Wc + h / 2 * k2W, Wpc + h / 2 * k2P);
            const double k4W = Wpc + h * k3P, k4P = dWp(xc + h, Wc + h * k3W, Wpc + h * k3P);
            Wc = Wc + (h / 6.0) * (k1W + 2 * k2W + 2 * k3W + k4W);
            Wpc = Wpc + (h / 6.0) * (k1P + 2 * k2P + 2 * k3P + k4P);
            xc += h;
            xs.push_back(xc);
            Ws.push_back(std::max(Wc, 0.0));
        }
        xt = xc;
        const size_t np = xs.size();                                              // :108-114
        r.resize(np);
        rhoShape.resize(np);
        for (size_t i = 0; i < np; ++i) {
            r[i] = xs[i] * r0;
            rhoShape[i] = OrcKingG(Ws[i]) / g0;
        }
        cumMassShape.assign(np, 0.0);                                             // :116-121
        for (size_t i = 1; i < np; ++i) {
            const double f0 = rhoShape[i - 1] * r[i - 1] * r[i - 1];
            const double f1 = rhoShape[i] * r[i] * r[i];
            cumMassShape[i] = cumMassShape[i - 1] + 0.5 * (f0 + f1) * (r[i] - r[i - 1]);
        }
        const double totalShape = cumMassShape.back();
        std::vector<double> massPhys(np), tail(np, 0.0);                          // :126-135
        for (size_t i = 0; i < np; ++i) massPhys[i] = totalMass * cumMassShape[i] / totalShape;
        for (size_t i = np - 1; i-- > 0;) {
            const double ra = r[i], rb = r[i + 1];
            const double f0 = (ra > 1e-300) ? rhoShape[i] * G * massPhys[i] / (ra * ra) : 0.0;
            const double f1 = rhoShape[i + 1] * G * massPhys[i + 1] / (rb * rb);
            tail[i] = tail[i + 1] + 0.5 * (f0 + f1) * (rb - ra);
        }
        sigma2.resize(np);                                                        // :136-139
        for (size_t i = 0; i < np; ++i) sigma2[i] = (rhoShape[i] > 1e-12) ? (tail[i] / rhoShape[i]) : 0.0;
    }
    // SampleRadius (King.hpp:144-153)
    double SampleRadius(double x1) const {
        const double target = x1 * cumMassShape.back();
        size_t i1 = static_cast<size_t>(std::lower_bound(cumMassShape.begin(), cumMassShape.end(), target) -
                                        cumMassShape.begin());
        i1 = std::min(std::max(i1, size_t{1}), cumMassShape.size() - 1);
        const size_t i0 = i1 - 1;
        const double c0 = cumMassShape[i0], c1 = cumMassShape[i1];
        const double frac = (c1 > c0) ? (target - c0) / (c1 - c0) : 0.0;
        return r[i0] * (1.0 - frac) + r[i1] * frac;
    }
    // Sigma2At (King.hpp:155-163)
    double Sigma2At(double rr) const {
        size_t i1 = static_cast<size_t>(std::lower_bound(r.begin(), r.end(), rr) - r.begin());
        i1 = std::min(std::max(i1, size_t{1}), r.size() - 1);
        const size_t i0 = i1 - 1;
        const double frac = (r[i1] > r[i0]) ? (rr - r[i0]) / (r[i1] - r[i0]) : 0.0;
        return sigma2[i0] * (1.0 - frac) + sigma2[i1] * frac;
    }
};
}  // namespace

// Dimensionless tidal radius x_t of the King(W0) profile (KingProfile::xt, King.hpp:106).
double orc_king_tidal_radius(double w0) { return OrcKingProfile(w0, 1.0, 1.0, 1.0).xt; }

// ic::King<3> (King.hpp:168-211).  Velocity components in the GCC (right-to-left) order of
// Vec<3> v(gg(rng), gg(rng), gg(rng)) (King.hpp:190), as in orc_ic_hernquist.
void orc_ic_king(int n, uint32_t seed, double totalMass, double coreRadius, double w0, double G, double* x,
                 double* y, double* z, double* vx, double* vy, double* vz, double* m) {
    const OrcKingProfile profile(w0, coreRadius, G, totalMass);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double mi = totalMass / n;
    for (int i = 0; i < n; ++i) {
        const double r = profile.SampleRadius(u(rng));                  // :180
        const double cosT = 2.0 * u(rng) - 1.0;                         // :182-186
        const double sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
        const double phi = 2.0 * M_PI * u(rng);
        x[i] = sinT * std::cos(phi) * r;
        y[i] = sinT * std::sin(phi) * r;
        z[i] = cosT * r;
        const double sigma = std::sqrt(std::max(0.0, profile.Sigma2At(r)));  // :188-191
        std::normal_distribution<double> gg(0.0, sigma);
        const double drawZ = gg(rng);
        const double drawY = gg(rng);
        const double drawX = gg(rng);
        vx[i] = drawX;
        vy[i] = drawY;
        vz[i] = drawZ;
        m[i] = mi;
    }
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0}, tm = 0.0;              // :195-209
    for (int i = 0; i < n; ++i) {
        cp[0] += m[i] * x[i];
        cp[1] += m[i] * y[i];
        cp[2] += m[i] * z[i];
        cv[0] += m[i] * vx[i];
        cv[1] += m[i] * vy[i];
        cv[2] += m[i] * vz[i];
        tm += m[i];
    }
    for (int k = 0; k < 3; ++k) {
        cp[k] /= tm;
        cv[k] /= tm;
    }
    for (int i = 0; i < n; ++i) {
        x[i] -= cp[0];
        y[i] -= cp[1];
        z[i] -= cp[2];
        vx[i] -= cv[0];
        vy[i] -= cv[1];
        vz[i] -= cv[2];
    }
}

// n successive draws of std::uniform_real_distribution<double>(lo, hi) on mt19937(seed):
// reproduces the inputs of the reference's own seeded self-tests.
void orc_uniform_sequence(uint32_t seed, double lo, double hi, int n, double* out) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(lo, hi);
    for (int i = 0; i < n; ++i) out[i] = u(rng);
}

// The 10000th output of a default-seeded mt19937 (C++ standard [rand.predef]).
uint32_t orc_mt19937_10000th(void) {
    std::mt19937 g;
    g.discard(9999);
    return static_cast<uint32_t>(g());
}

// ---------------------------------------------------------------- tree
void* orc_tree_build(int n, const double* x, const double* y, const double* z, const double* m, int ncrit,
                     int maxDepth, int withQuad) {
    Tree* t = new Tree();
    t->ncrit = ncrit;
    t->maxDepth = maxDepth;
    const Pts p{x, y, z, m, n};
    double c[3], h;
    RootBox(p, c, &h);
    NewNode(*t, c[0], c[1], c[2], h, -1, 0);
    for (int i = 0; i < n; ++i) Insert(*t, p, 0, i, 0);
    Finalize(*t, p);
    if (withQuad) Quadrupoles(*t, p);
    return t;
}

void orc_tree_free(void* h) { delete static_cast<Tree*>(h); }

int orc_tree_num_nodes(const void* h) { return static_cast<int>(static_cast<const Tree*>(h)->mass.size()); }

// Export in canonical numbering (DFS pre-order, children in ascending octant),
// which is the order Finalize's DFS visits nodes (Tree.cpp:157-175).
// Any output pointer may be NULL.  geom: [cx,cy,cz,half]; mom: [comx,comy,comz,mass,maxR];
// quad: [Qxx,Qyy,Qxy,Qxz,Qyz]; topo: [parent, first, count, depth, childCount];
// children: 8 per node (-1 = empty).
void orc_tree_export(const void* h, double* geom, double* mom, double* quad, int* topo, int* children,
                     int* order) {
    const Tree& t = *static_cast<const Tree*>(h);
    const int nn = static_cast<int>(t.mass.size());
    for (int ci = 0; ci < nn; ++ci) {
        const int nd = t.fromCanon[ci];
        if (geom) {
            geom[4 * ci + 0] = t.cx[nd];
            geom[4 * ci + 1] = t.cy[nd];
            geom[4 * ci + 2] = t.cz[nd];
            geom[4 * ci + 3] = t.half[nd];
        }
        if (mom) {
            mom[5 * ci + 0] = t.comx[nd];
            mom[5 * ci + 1] = t.comy[nd];
            mom[5 * ci + 2] = t.comz[nd];
            mom[5 * ci + 3] = t.mass[nd];
            mom[5 * ci + 4] = t.maxR[nd];
        }
        if (quad) {
            quad[5 * ci + 0] = t.hasQ ? t.Qxx[nd] : 0.0;
            quad[5 * ci + 1] = t.hasQ ? t.Qyy[nd] : 0.0;
            quad[5 * ci + 2] = t.hasQ ? t.Qxy[nd] : 0.0;
            quad[5 * ci + 3] = t.hasQ ? t.Qxz[nd] : 0.0;
            quad[5 * ci + 4] = t.hasQ ? t.Qyz[nd] : 0.0;
        }
        if (topo) {
            topo[5 * ci + 0] = t.parent[nd] < 0 ? -1 : t.canon[t.parent[nd]];
            topo[5 * ci + 1] = t.first[nd];
            topo[5 * ci + 2] = t.count[nd];
            topo[5 * ci + 3] = t.depth[nd];
            topo[5 * ci + 4] = t.childCount[nd];
        }
        if (children) {
            for (int c = 0; c < 8; ++c)
                children[8 * ci + c] = t.child[nd][c] < 0 ? -1 : t.canon[t.child[nd][c]];
        }
    }
    if (order) std::memcpy(order, t.order.data(), sizeof(int) * t.order.size());
}

// ---------------------------------------------------------------- forces
// ComputeAccelBarnesHutGroup (Solvers.cpp:279-355) on a tree built by orc_tree_build
// from the same arrays.  nFar / nNear (nullable) receive, per canonical node id, the
// far-node count and the near P2P source count (near leaves + own leaf, self excluded)
// of that leaf's group; zero for internal nodes.
void orc_accel_bh(const void* h, int n, const double* x, const double* y, const double* z, const double* m,
                  double G, double eps2_, double theta, int spline, double* ax, double* ay, double* az,
                  int64_t* nFar, int64_t* nNear) {
    const Tree& t = *static_cast<const Tree*>(h);
    const Pts p{x, y, z, m, n};
    const Soft eps2 = MakeSoft(eps2_, spline);
    const double theta2 = theta * theta;
    std::vector<int> leaves;
    for (int nd = 0; nd < static_cast<int>(t.mass.size()); ++nd)
        if (t.childCount[nd] == 0 && t.count[nd] > 0) leaves.push_back(nd);
    if (nFar) std::memset(nFar, 0, sizeof(int64_t) * t.mass.size());
    if (nNear) std::memset(nNear, 0, sizeof(int64_t) * t.mass.size());
    const int nl = static_cast<int>(leaves.size());
#pragma omp parallel
    {
        std::vector<int> far, near;
#pragma omp for schedule(dynamic, 32)
        for (int li = 0; li < nl; ++li) {
            const int g = leaves[li];
            far.clear();
            near.clear();
            Collect(t, 0, g, t.comx[g], t.comy[g], t.comz[g], t.maxR[g], theta2, far, near);
            int64_t nearParticles = t.count[g] - 1;
            for (int lf : near) nearParticles += t.count[lf];
            if (nFar) nFar[t.canon[g]] = static_cast<int64_t>(far.size());
            if (nNear) nNear[t.canon[g]] = nearParticles;
            for (int k = t.first[g]; k < t.first[g] + t.count[g]; ++k) {
                const int i = t.order[k];
                const double xi = x[i], yi = y[i], zi = z[i];
                double sx = 0.0, sy = 0.0, sz = 0.0;
                for (int nd : far) {
                    const double dx = t.comx[nd] - xi, dy = t.comy[nd] - yi, dz = t.comz[nd] - zi;
                    const double dist2 = dx * dx + dy * dy + dz * dz;
                    AccumPair(dx, dy, dz, G * t.mass[nd], eps2, sx, sy, sz);
                    AddQuad(t, nd, dx, dy, dz, dist2, G, sx, sy, sz);
                }
                for (int lf : near) {
                    for (int k2 = t.first[lf]; k2 < t.first[lf] + t.count[lf]; ++k2) {
                        const int q = t.order[k2];
                        AccumPair(x[q] - xi, y[q] - yi, z[q] - zi, G * m[q], eps2, sx, sy, sz);
                    }
                }
                for (int k2 = t.first[g]; k2 < t.first[g] + t.count[g]; ++k2) {
                    if (k2 == k) continue;
                    const int q = t.order[k2];
                    AccumPair(x[q] - xi, y[q] - yi, z[q] - zi, G * m[q], eps2, sx, sy, sz);
                }
                ax[i] = sx;
                ay[i] = sy;
                az[i] = sz;
            }
        }
    }
    (void)p;
}

// The interaction lists CollectGroupInteractions (Solvers.cpp:248-277) builds for every
// leaf, as CSR arrays over canonical ids: leaf k (in pre-order among leaves) has far
// cells far_idx[far_ptr[k] .. far_ptr[k+1]) and near leaves near_idx[...] in walk order.
// Call with far_idx == NULL to get the totals in far_ptr[nleaves], near_ptr[nleaves].
void orc_group_lists(const void* h, double theta, int64_t* far_ptr, int32_t* far_idx, int64_t* near_ptr,
                     int32_t* near_idx) {
    const Tree& t = *static_cast<const Tree*>(h);
    const double theta2 = theta * theta;
    std::vector<int> far, near;
    int64_t nf = 0, nq = 0;
    int k = 0;
    for (int ci = 0; ci < static_cast<int>(t.fromCanon.size()); ++ci) {
        const int g = t.fromCanon[ci];
        if (t.childCount[g] != 0 || t.count[g] == 0) continue;
        far.clear();
        near.clear();
        Collect(t, 0, g, t.comx[g], t.comy[g], t.comz[g], t.maxR[g], theta2, far, near);
        far_ptr[k] = nf;
        near_ptr[k] = nq;
        if (far_idx)
            for (int nd : far) far_idx[nf++] = t.canon[nd];
        else
            nf += static_cast<int64_t>(far.size());
        if (near_idx)
            for (int nd : near) near_idx[nq++] = t.canon[nd];
        else
            nq += static_cast<int64_t>(near.size());
        ++k;
    }
    far_ptr[k] = nf;
    near_ptr[k] = nq;
}

// ComputeAccelBarnesHutPerParticle (Solvers.cpp:359-382): relative MAC when
// relative != 0 and aOld (nullable) holds each particle's previous |a|.
// counts (nullable, 2 entries) receives the totals of M2P and P2P interactions.
void orc_accel_bh_pp(const void* h, int n, const double* x, const double* y, const double* z, const double* m,
                     double G, double eps2_, double theta, int spline, int relative, double alpha, const double* aOld,
                     double* ax, double* ay, double* az, int64_t* counts) {
    const Tree& t = *static_cast<const Tree*>(h);
    const Pts p{x, y, z, m, n};
    const Soft eps2 = MakeSoft(eps2_, spline);
    const double theta2 = theta * theta;
    int64_t nm2p = 0, np2p = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nm2p, np2p)
    for (int i = 0; i < n; ++i) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        int64_t cnt[2] = {0, 0};
        WalkPP(t, p, 0, i, x[i], y[i], z[i], G, theta2, eps2, relative, alpha, aOld ? aOld[i] : 0.0, sx, sy, sz,
               cnt);
        ax[i] = sx;
        ay[i] = sy;
        az[i] = sz;
        nm2p += cnt[0];
        np2p += cnt[1];
    }
    if (counts) {
        counts[0] = nm2p;
        counts[1] = np2p;
    }
}

// ComputeAccelBarnesHutTargets (Solvers.cpp:384-405): the per-particle geometric walk onto
// targets[0..nt) (outputs by list slot).
void orc_accel_bh_targets(const void* h, int n, const double* x, const double* y, const double* z, const double* m,
                          double G, double eps2_, double theta, int spline, int nt, const int* targets, double* ax,
                          double* ay, double* az) {
    const Tree& t = *static_cast<const Tree*>(h);
    const Pts p{x, y, z, m, n};
    const Soft eps2 = MakeSoft(eps2_, spline);
    const double theta2 = theta * theta;
#pragma omp parallel for schedule(dynamic, 64)
    for (int k = 0; k < nt; ++k) {
        const int i = targets[k];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        WalkPP(t, p, 0, i, x[i], y[i], z[i], G, theta2, eps2, 0, 0.0, 0.0, sx, sy, sz);
        ax[k] = sx;
        ay[k] = sy;
        az[k] = sz;
    }
}

// ComputeAccelDirect (Solvers.cpp:77-106).
void orc_accel_direct(int n, const double* x, const double* y, const double* z, const double* m, double G,
                      double eps2_, int spline, double* ax, double* ay, double* az) {
    const Soft eps2 = MakeSoft(eps2_, spline);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const double xi = x[i], yi = y[i], zi = z[i];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            AccumPair(x[j] - xi, y[j] - yi, z[j] - zi, G * m[j], eps2, sx, sy, sz);
        }
        ax[i] = sx;
        ay[i] = sy;
        az[i] = sz;
    }
}

// Direct sum onto a subset of targets (ComputeAccelDirectTargets, Solvers.cpp:108-134).
void orc_accel_direct_targets(int n, const double* x, const double* y, const double* z, const double* m, double G,
                              double eps2_, int spline, int nt, const int* targets, double* ax, double* ay,
                              double* az) {
    const Soft eps2 = MakeSoft(eps2_, spline);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; ++t) {
        const int i = targets[t];
        const double xi = x[i], yi = y[i], zi = z[i];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            AccumPair(x[j] - xi, y[j] - yi, z[j] - zi, G * m[j], eps2, sx, sy, sz);
        }
        ax[t] = sx;
        ay[t] = sy;
        az[t] = sz;
    }
}

// ---------------------------------------------------------------- integrator
namespace {

// System::ComputeAccel for the native solvers (System.hpp:43-78): Direct, or Barnes-Hut on
// a tree rebuilt for this evaluation (quadrupoles when withQuad), grouped walk for the
// geometric MAC, per-particle walk with aOld for the relative MAC (Solvers.cpp:407-414).
void ComputeAccel(int n, const double* x, const double* y, const double* z, const double* m, double* ax,
                  double* ay, double* az, double G, double eps2, int spline, double theta, int solver, int withQuad,
                  int relative, double alpha, const double* aOld) {
    if (solver == 0) {
        orc_accel_direct(n, x, y, z, m, G, eps2, spline, ax, ay, az);
        return;
    }
    void* t = orc_tree_build(n, x, y, z, m, 8, 40, withQuad);
    if (relative)
        orc_accel_bh_pp(t, n, x, y, z, m, G, eps2, theta, spline, 1, alpha, aOld, ax, ay, az, nullptr);
    else
        orc_accel_bh(t, n, x, y, z, m, G, eps2, theta, spline, ax, ay, az, nullptr, nullptr);
    orc_tree_free(t);
}

// System::CaptureAOld (System.hpp:302-310).
void CaptureAOld(int n, const double* ax, const double* ay, const double* az, double* aOld) {
    for (int i = 0; i < n; ++i) aOld[i] = std::sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
}

}  // namespace

// ChooseAdaptiveDt (Integrator.cpp:41-59).
double orc_adaptive_dt(int n, const double* ax, const double* ay, const double* az, double eta, double eps,
                       double dtFallback) {
    if (n == 0 || eps <= 0.0) return dtFallback;
    double minTau = std::numeric_limits<double>::max();
    for (int i = 0; i < n; ++i) {
        const double a = std::sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
        if (a > 0.0) minTau = std::min(minTau, std::sqrt(eps / a));
    }
    if (minTau == std::numeric_limits<double>::max()) return dtFallback;
    const double dt = eta * minTau;
    return std::clamp(dt, dtFallback / 4096.0, dtFallback);
}

// LeapfrogStep (Integrator.cpp:9-39) as System::Step drives it (System.hpp:89-108): global
// steps, optional adaptive dt (from a(t)), aOld captured after every step for the relative
// MAC.  a[] must hold a(t) on entry and aOld (nullable unless relative) the |a| captured
// with it; both are updated.  Returns the time advanced.
double orc_steps(int n, double* x, double* y, double* z, double* vx, double* vy, double* vz, const double* m,
                 double* ax, double* ay, double* az, double* aOld, double G, double eps2, int spline, double theta,
                 int solver, int withQuad, int relative, double alpha, int adaptive, double eta, double dt0,
                 int nsteps) {
    double t = 0.0;
    std::vector<double> sx(n), sy(n), sz(n);
    const double eps = std::sqrt(eps2);
    for (int s = 0; s < nsteps; ++s) {
        const double dt = adaptive ? orc_adaptive_dt(n, ax, ay, az, eta, eps, dt0) : dt0;
        const double hdt = 0.5 * dt;
        for (int i = 0; i < n; ++i) {
            vx[i] += hdt * ax[i];
            vy[i] += hdt * ay[i];
            vz[i] += hdt * az[i];
        }
        for (int i = 0; i < n; ++i) {
            x[i] += dt * vx[i];
            y[i] += dt * vy[i];
            z[i] += dt * vz[i];
        }
        ComputeAccel(n, x, y, z, m, sx.data(), sy.data(), sz.data(), G, eps2, spline, theta, solver, withQuad,
                     relative, alpha, aOld);
        std::copy(sx.begin(), sx.end(), ax);
        std::copy(sy.begin(), sy.end(), ay);
        std::copy(sz.begin(), sz.end(), az);
        for (int i = 0; i < n; ++i) {
            vx[i] += hdt * ax[i];
            vy[i] += hdt * ay[i];
            vz[i] += hdt * az[i];
        }
        if (aOld) CaptureAOld(n, ax, ay, az, aOld);
        t += dt;
    }
    return t;
}

// The Plummer, geometric, fixed-dt special case of orc_steps.
void orc_leapfrog(int n, double* x, double* y, double* z, double* vx, double* vy, double* vz, const double* m,
                  double* ax, double* ay, double* az, double G, double eps2, double theta, int solver, int withQuad,
                  double dt, int nsteps) {
    orc_steps(n, x, y, z, vx, vy, vz, m, ax, ay, az, nullptr, G, eps2, 0, theta, solver, withQuad, 0, 0.0, 0, 0.0, dt,
              nsteps);
}

// Rungs of System::BlockStep (System.hpp:249-265) from a; returns the deepest rung.
int orc_rungs(int n, const double* ax, const double* ay, const double* az, double eta, double eps, double base,
              int maxRung, int* rung) {
    maxRung = std::max(0, maxRung);
    int rMax = 0;
    for (int i = 0; i < n; ++i) {
        const double amag = std::sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
        int k = 0;
        if (amag > 0.0 && eps > 0.0) {
            const double tau = eta * std::sqrt(eps / amag);
            if (tau < base) k = std::min(maxRung, static_cast<int>(std::ceil(std::log2(base / tau))));
        }
        rung[i] = k;
        rMax = std::max(rMax, k);
    }
    return rMax;
}

// nsteps calls of System::BlockStep (System.hpp:239-294) with the Direct or Barnes-Hut
// solver: the closing set of every substep is evaluated on a fresh tree by the per-particle
// geometric walk (ComputeAccelTargets, System.hpp:220-230) or by the direct sum onto those
// targets.  rung[] receives the last coarse step's rungs; *evals accumulates the
// evaluated targets (System::ParticleEvals).  Returns the deepest rung of the last step.
int orc_block_steps(int n, double* x, double* y, double* z, double* vx, double* vy, double* vz, const double* m,
                    double* ax, double* ay, double* az, int* rung, double G, double eps2, int spline, double theta,
                    int solver, int withQuad, double eta, double base, int maxRung, int nsteps, int64_t* evals) {
    const double eps = std::sqrt(eps2);
    int rMax = 0;
    std::vector<int> active;
    std::vector<double> tx, ty, tz;
    auto half_kick = [&](int i, double h) {
        vx[i] += h * ax[i];
        vy[i] += h * ay[i];
        vz[i] += h * az[i];
    };
    for (int s = 0; s < nsteps; ++s) {
        rMax = orc_rungs(n, ax, ay, az, eta, eps, base, maxRung, rung);
        const int nSub = 1 << rMax;
        const double dtMin = base / nSub;
        for (int sub = 0; sub < nSub; ++sub) {
            for (int i = 0; i < n; ++i) {
                const int stepUnits = 1 << (rMax - rung[i]);
                if (sub % stepUnits == 0) half_kick(i, 0.5 * (base / (1 << rung[i])));
            }
            for (int i = 0; i < n; ++i) {
                x[i] += dtMin * vx[i];
                y[i] += dtMin * vy[i];
                z[i] += dtMin * vz[i];
            }
            active.clear();
            for (int i = 0; i < n; ++i) {
                const int stepUnits = 1 << (rMax - rung[i]);
                if ((sub + 1) % stepUnits == 0) active.push_back(i);
            }
            const int nt = static_cast<int>(active.size());
            if (nt > 0) {
                if (evals) *evals += nt;
                tx.assign(nt, 0.0);
                ty.assign(nt, 0.0);
                tz.assign(nt, 0.0);
                if (solver == 0) {
                    orc_accel_direct_targets(n, x, y, z, m, G, eps2, spline, nt, active.data(), tx.data(), ty.data(),
                                             tz.data());
                } else {
                    void* t = orc_tree_build(n, x, y, z, m, 8, 40, withQuad);
                    orc_accel_bh_targets(t, n, x, y, z, m, G, eps2, theta, spline, nt, active.data(), tx.data(),
                                         ty.data(), tz.data());
                    orc_tree_free(t);
                }
                for (int k = 0; k < nt; ++k) {
                    ax[active[k]] = tx[k];
                    ay[active[k]] = ty[k];
                    az[active[k]] = tz[k];
                }
            }
            for (int i : active) half_kick(i, 0.5 * (base / (1 << rung[i])));
        }
    }
    return rMax;
}

// ---------------------------------------------------------------- diagnostics
// System::TotalEnergy (System.hpp:119-164).  Per-i partial sums are added in i order so
// the result does not depend on the thread count.
double orc_total_energy(int n, const double* x, const double* y, const double* z, const double* vx,
                        const double* vy, const double* vz, const double* m, double G, double eps2, int spline) {
    const Soft soft = MakeSoft(eps2, spline);
    double kinetic = 0.0;
    for (int i = 0; i < n; ++i) {
        const double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        kinetic += 0.5 * m[i] * v2;
    }
    std::vector<double> local(static_cast<size_t>(n), 0.0);
#pragma omp parallel for schedule(dynamic, 32)
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = i + 1; j < n; ++j) {
            const double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            const double mij = m[i] * m[j];
            if (soft.spline) {
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                const double q = (r <= 1e-300) ? 0.0 : r / soft.eps;
                acc += G * mij * SplinePotential3D(q) / soft.eps;
            } else {
                const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                acc -= G * mij / std::sqrt(r2);
            }
        }
        local[i] = acc;
    }
    double potential = 0.0;
    for (int i = 0; i < n; ++i) potential += local[i];
    return kinetic + potential;
}

// CenterOfMass / AngularMomentum (System.hpp:166-193); momentum as DeckSim.cpp:147-153.
void orc_moments(int n, const double* x, const double* y, const double* z, const double* vx, const double* vy,
                 const double* vz, const double* m, double* com3, double* L3, double* P3) {
    double tm = 0.0, c[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        c[0] += m[i] * x[i];
        c[1] += m[i] * y[i];
        c[2] += m[i] * z[i];
        tm += m[i];
    }
    if (tm > 0.0)
        for (int k = 0; k < 3; ++k) c[k] /= tm;
    double L[3] = {0, 0, 0}, P[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const double rx = x[i] - c[0], ry = y[i] - c[1], rz = z[i] - c[2];
        L[0] += m[i] * (ry * vz[i] - rz * vy[i]);
        L[1] += m[i] * (rz * vx[i] - rx * vz[i]);
        L[2] += m[i] * (rx * vy[i] - ry * vx[i]);
        P[0] += m[i] * vx[i];
        P[1] += m[i] * vy[i];
        P[2] += m[i] * vz[i];
    }
    for (int k = 0; k < 3; ++k) {
        if (com3) com3[k] = c[k];
        if (L3) L3[k] = L[k];
        if (P3) P3[k] = P[k];
    }
}

// ---------------------------------------------------------------- derived key contract
// Root box exactly as Tree.cpp:54-88 builds it.
void orc_root_box(int n, const double* x, const double* y, const double* z, double* center3, double* half) {
    const Pts p{x, y, z, nullptr, n};
    RootBox(p, center3, half);
}

// Octant digits of each particle for levels 1..levels (levels <= 40), MSB first, from
// the FP64 centre recurrence of Tree.cpp:96-99 / 11-20.  Digits of levels 1..21 go to
// keyHi (bits 62..0), levels 22..40 to keyLo (bits 56..0); keyLo may be NULL.
void orc_morton_keys(int n, const double* x, const double* y, const double* z, const double* center3,
                     double half, int levels, uint64_t* keyHi, uint64_t* keyLo) {
    for (int i = 0; i < n; ++i) {
        double cx = center3[0], cy = center3[1], cz = center3[2], h = half;
        uint64_t hi = 0, lo = 0;
        for (int l = 1; l <= levels; ++l) {
            const int oct = OctantOf(cx, cy, cz, x[i], y[i], z[i]);
            h = 0.5 * h;
            cx = cx + ((oct & 1) ? h : -h);
            cy = cy + ((oct & 2) ? h : -h);
            cz = cz + ((oct & 4) ? h : -h);
            if (l <= 21)
                hi = (hi << 3) | static_cast<uint64_t>(oct);
            else
                lo = (lo << 3) | static_cast<uint64_t>(oct);
        }
        if (levels < 21) hi <<= 3 * (21 - levels);
        if (levels > 21 && levels < 40) lo <<= 3 * (40 - levels);
        keyHi[i] = hi;
        if (keyLo) keyLo[i] = lo;
    }
}

} // extern "C"
