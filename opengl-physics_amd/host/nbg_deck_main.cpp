// nbg_deck -- command line of the GPU backend; the modes of
// projects/02_nbody_gravity/src/main.cpp on the gravity path:
//   --deck <file> --out <dir> [--frames N]                     (main.cpp:397-527)
//   --bench <solver> [--bench-n N] [--bench-theta T] [--bench-reps R] [--bench-eps E]
//           [--bench-order K]                                  (main.cpp:282-393)
//   --dump-ic <path> [--bench-n N]                             (main.cpp:249-280)
// with the reference's defaults (main.cpp:52-65).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ngrav_host.h"

int main(int argc, char** argv) {
    std::string deck, out = "out", bench, dump;
    int frames = 0, quiet = 0;
    int bench_n = 10000, bench_reps = 5, bench_order = 5;
    double bench_theta = 0.5, bench_eps = 0.02;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--deck" && i + 1 < argc) deck = argv[++i];
        else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--frames" && i + 1 < argc) frames = std::atoi(argv[++i]);
        else if (a == "--quiet") quiet = 1;
        else if (a == "--bench" && i + 1 < argc) bench = argv[++i];
        else if (a == "--bench-n" && i + 1 < argc) bench_n = std::atoi(argv[++i]);
        else if (a == "--bench-theta" && i + 1 < argc) bench_theta = std::atof(argv[++i]);
        else if (a == "--bench-reps" && i + 1 < argc) bench_reps = std::atoi(argv[++i]);
        else if (a == "--bench-order" && i + 1 < argc) bench_order = std::atoi(argv[++i]);
        else if (a == "--bench-eps" && i + 1 < argc) bench_eps = std::atof(argv[++i]);
        else if (a == "--dump-ic" && i + 1 < argc) dump = argv[++i];
        else {
            std::fprintf(stderr,
                         "usage: nbg_deck --deck <file.toml> [--out <dir>] [--frames N] [--quiet]\n"
                         "       nbg_deck --bench <solver> [--bench-n N] [--bench-theta T] [--bench-reps R] "
                         "[--bench-eps E] [--bench-order K]\n"
                         "       nbg_deck --dump-ic <path> [--bench-n N]\n");
            return 2;
        }
    }
    if (!dump.empty()) {
        if (ngh_dump_ic(dump.c_str(), bench_n) != 0) {
            std::fprintf(stderr, "dump-ic: cannot write '%s'\n", dump.c_str());
            return 1;
        }
        std::fprintf(stderr, "# dump-ic path=%s N=%d eps=%.5g G=%.4f seed=%u format=raw-f32-x,y,z,m,vx,vy,vz\n",
                     dump.c_str(), bench_n, bench_eps, 1.0, 1u);
        return 0;
    }
    if (!bench.empty()) {
        char csv[256] = {0}, err[512] = {0};
        if (ngh_bench(bench.c_str(), bench_n, bench_theta, bench_eps, bench_reps, bench_order, csv, sizeof(csv), err,
                      sizeof(err)) != 0) {
            std::fprintf(stderr, "nbg_deck: %s\n", err);
            return 2;
        }
        std::fprintf(stderr, "solver,n,theta,order,ms_per_eval,mean_rel_err\n");
        std::printf("%s\n", csv);
        return 0;
    }
    if (deck.empty()) {
        std::fprintf(stderr, "nbg_deck: --deck, --bench or --dump-ic is required\n");
        return 2;
    }
    char err[512] = {0};
    if (ngh_run_deck(deck.c_str(), out.c_str(), frames, quiet, err, sizeof(err)) != 0) {
        std::fprintf(stderr, "nbg_deck: %s\n", err);
        return 1;
    }
    return 0;
}
