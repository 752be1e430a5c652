/*
 * ngrav_host.h -- host-side C entry points of libngrav_host.so (C++ host code above the
 * nbg C ABI).  Initial-condition samplers: FP64 arrays of length n, caller-owned.
 * Return 0 on success, -1 on a bad argument.
 */
#ifndef NGRAV_HOST_H
#define NGRAV_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ic::Plummer<3> (projects/nbody_core/include/ngrav/ic/Plummer.hpp:26-112). */
int ngh_ic_plummer(int n, uint32_t seed, double total_mass, double scale_radius, double G, double* x, double* y,
                   double* z, double* vx, double* vy, double* vz, double* m);

/* ic::Hernquist<3> (projects/nbody_core/include/ngrav/ic/Hernquist.hpp:131-181). */
int ngh_ic_hernquist(int n, uint32_t seed, double total_mass, double scale_radius, double G, double* x, double* y,
                     double* z, double* vx, double* vy, double* vz, double* m);

/* ic::King<3> (projects/nbody_core/include/ngrav/ic/King.hpp:168-211). */
int ngh_ic_king(int n, uint32_t seed, double total_mass, double core_radius, double w0, double G, double* x,
                double* y, double* z, double* vx, double* vy, double* vz, double* m);

/* Uniform-sphere cluster (Cluster<3>, projects/nbody_core/src/Scenarios.cpp:104-142): R = 1,
 * M = 1, Gaussian velocities of dispersion 0.3; cold != 0 zeroes the velocities after the
 * same draws (ColdUniform, Scenarios.cpp:201-214, the cold_collapse scenario). */
int ngh_ic_cluster(int n, uint32_t seed, int cold, double* x, double* y, double* z, double* vx, double* vy,
                   double* vz, double* m);

/* Batched-ensemble layout (no reference counterpart: the reference runs one System).
 * N_k = round(exp(U[ln nmin, ln nmax])) with U drawn from std::mt19937(seed) through
 * std::uniform_real_distribution<double>, k = 0..nclusters-1; kinds[k] = k % 2 (0 Plummer,
 * 1 King).  kinds may be NULL.  Returns the total particle count, or -1 on a bad argument. */
int64_t ngh_ensemble_sizes(int32_t nclusters, uint32_t seed, int64_t nmin, int64_t nmax, int64_t* sizes,
                           int32_t* kinds);

/* Deck-driven batch run on the GPU backend: DeckSim::Configure + fw::RunHeadless +
 * fw::OutputWriter (projects/nbody_core/src/DeckSim.cpp:26-169,
 * framework/src/HeadlessRunner.cpp:11-41, framework/src/OutputWriter.cpp:38-169).
 * frames <= 0 uses the deck's time.frames.  Returns 0, or -1 with a message in err. */
int ngh_run_deck(const char* deck_path, const char* out_dir, int frames, int quiet, char* err, int err_len);

/* --dump-ic of the reference command line (projects/02_nbody_gravity/src/main.cpp:249-280):
 * the seed-1 Plummer sphere of N bodies (G = M = a = 1) as raw float32 records
 * x,y,z,m,vx,vy,vz -- the ic_file format.  Returns 0, or -1 on a bad argument or I/O error. */
int ngh_dump_ic(const char* path, int n);

/* --bench of the reference command line (main.cpp:282-393) on the GPU backend: mean wall
 * time of `reps` force evaluations (after one warm-up, upload and readback included) of
 * `solver` (direct | gpu_direct | barnes_hut | bh | gpu_bh) on the seed-1 Plummer sphere with
 * softening eps and opening angle theta, and the mean per-particle relative error against
 * the device's compensated direct sum (-1 above N = 60,000).  Writes the CSV line
 * "solver,n,theta,order,ms_per_eval,mean_rel_err" into csv.  Returns 0, or -1 with a
 * message in err. */
int ngh_bench(const char* solver, int n, double theta, double eps, int reps, int order, char* csv, int csv_len,
              char* err, int err_len);

#ifdef __cplusplus
}
#endif

#endif /* NGRAV_HOST_H */
