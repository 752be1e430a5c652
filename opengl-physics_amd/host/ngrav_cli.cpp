// ngrav_cli.cpp -- the measurement side of the reference's command line on the GPU backend:
//   --dump-ic  projects/02_nbody_gravity/src/main.cpp:249-280: the seed-1 Plummer sphere of
//              --bench written as raw float32 records x,y,z,m,vx,vy,vz (the ic_file format
//              of Scenarios.cpp:292-329);
//   --bench    main.cpp:282-393: one solver's force evaluation on that sphere, mean wall
//              time over `reps` evaluations after one warm-up, and the mean per-particle
//              relative error against a direct sum, printed as the reference's CSV line
//              `solver,n,theta,order,ms_per_eval,mean_rel_err`.
// The GPU solvers of the reference (gpu_direct, gpu_bh) and its CPU solver names
// (direct, barnes_hut / bh) all run on the device here: the direct sum and the grouped
// Barnes-Hut walk with quadrupoles.  Like the reference's GPU entry points, a timed
// evaluation includes the upload of positions and masses and the readback of the
// accelerations.  The error yardstick is the device's compensated FP32 direct sum
// (the reference uses its FP64 CPU direct sum); as in the reference it is skipped above
// N = 60,000 and the column is then -1.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbg.h"
#include "../../include/ngrav_host.h"

namespace {

constexpr double kBenchG = 1.0;        // main.cpp:244-247
constexpr uint32_t kBenchSeed = 1;

struct Sphere {
    std::vector<double> x, y, z, vx, vy, vz, m;
};

Sphere BenchSphere(int n) {
    Sphere s;
    for (auto* v : {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.m}) v->resize(static_cast<size_t>(n));
    if (ngh_ic_plummer(n, kBenchSeed, 1.0, 1.0, kBenchG, s.x.data(), s.y.data(), s.z.data(), s.vx.data(),
                       s.vy.data(), s.vz.data(), s.m.data()) != 0)
        throw std::runtime_error("ic::Plummer failed");
    return s;
}

void Fail(char* err, int len, const std::string& msg) {
    if (err && len > 0) std::snprintf(err, static_cast<size_t>(len), "%s", msg.c_str());
}

}  // namespace

extern "C" {

int ngh_dump_ic(const char* path, int n) {
    if (!path || n < 1) return -1;
    try {
        const Sphere s = BenchSphere(n);
        std::FILE* f = std::fopen(path, "wb");
        if (!f) return -1;
        float rec[7];
        for (int i = 0; i < n; ++i) {
            const size_t k = static_cast<size_t>(i);
            rec[0] = static_cast<float>(s.x[k]);
            rec[1] = static_cast<float>(s.y[k]);
            rec[2] = static_cast<float>(s.z[k]);
            rec[3] = static_cast<float>(s.m[k]);
            rec[4] = static_cast<float>(s.vx[k]);
            rec[5] = static_cast<float>(s.vy[k]);
            rec[6] = static_cast<float>(s.vz[k]);
            if (std::fwrite(rec, sizeof(float), 7, f) != 7) {
                std::fclose(f);
                return -1;
            }
        }
        return std::fclose(f) == 0 ? 0 : -1;
    } catch (const std::exception&) {
        return -1;
    }
}

int ngh_bench(const char* solver, int n, double theta, double eps, int reps, int order, char* csv, int csv_len,
              char* err, int err_len) {
    try {
        if (!solver || n < 1 || reps < 1) throw std::runtime_error("bench: bad arguments");
        const std::string name(solver);
        int kind;
        if (name == "direct" || name == "gpu_direct") kind = NBG_SOLVER_DIRECT;
        else if (name == "barnes_hut" || name == "bh" || name == "gpu_bh" || name == "gpu_barnes_hut")
            kind = NBG_SOLVER_BARNES_HUT;
        else
            throw std::runtime_error("bench: unknown solver '" + name +
                                     "' (direct|barnes_hut|gpu_direct|gpu_bh; the FMM solvers are not on the GPU "
                                     "backend)");
        const Sphere s = BenchSphere(n);
        std::vector<float> x(s.x.begin(), s.x.end()), y(s.y.begin(), s.y.end()), z(s.z.begin(), s.z.end()),
            m(s.m.begin(), s.m.end());
        nbg_ctx* ctx = nbg_create(0, n);
        if (!ctx) throw std::runtime_error("nbg_create failed");
        struct Closer {
            nbg_ctx* c;
            ~Closer() { nbg_destroy(c); }
        } closer{ctx};
        auto chk = [&](int rc, const char* what) {
            if (rc != NBG_OK) throw std::runtime_error(std::string(what) + ": " + nbg_last_error(ctx));
        };
        nbg_params p{};
        p.solver = kind;
        p.multipole = NBG_QUADRUPOLE;
        p.G = kBenchG;
        p.eps2 = eps * eps;
        p.theta = theta;
        p.dt = 1e-3;
        p.flags = NBG_FLAG_KAHAN;
        std::vector<float> ax(static_cast<size_t>(n)), ay(ax), az(ax);
        auto eval = [&]() {
            chk(nbg_upload(ctx, n, x.data(), y.data(), z.data(), m.data(), nullptr, nullptr, nullptr), "nbg_upload");
            chk(nbg_accel(ctx, &p, nullptr), "nbg_accel");
            chk(nbg_download(ctx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ax.data(), ay.data(),
                             az.data()),
                "nbg_download");
        };
        eval();  // warm-up
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r) eval();
        const auto t1 = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / reps;
        double mean_rel = -1.0;
        if (n <= 60000) {
            nbg_params pd = p;
            pd.solver = NBG_SOLVER_DIRECT;
            std::vector<float> rx(ax.size()), ry(ax.size()), rz(ax.size());
            chk(nbg_accel(ctx, &pd, nullptr), "nbg_accel");
            chk(nbg_download(ctx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, rx.data(), ry.data(),
                             rz.data()),
                "nbg_download");
            double sum = 0.0;
            for (int i = 0; i < n; ++i) {
                const size_t k = static_cast<size_t>(i);
                const double ex = ax[k] - static_cast<double>(rx[k]), ey = ay[k] - static_cast<double>(ry[k]),
                             ez = az[k] - static_cast<double>(rz[k]);
                const double rr = std::sqrt(static_cast<double>(rx[k]) * rx[k] + static_cast<double>(ry[k]) * ry[k] +
                                            static_cast<double>(rz[k]) * rz[k]);
                sum += std::sqrt(ex * ex + ey * ey + ez * ez) / std::max(rr, 1e-30);
            }
            mean_rel = sum / n;
        }
        if (csv && csv_len > 0)
            std::snprintf(csv, static_cast<size_t>(csv_len), "%s,%d,%.3f,%d,%.4f,%.6e", name.c_str(), n, theta, order,
                          ms, mean_rel);
        return 0;
    } catch (const std::exception& e) {
        Fail(err, err_len, e.what());
        return -1;
    }
}

}  // extern "C"
