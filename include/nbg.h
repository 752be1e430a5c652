/*
 * nbg.h -- C ABI of the MI355X gravity library (libnbg_hip.so).
 *
 * Plain C: pointers, sizes and POD structs only; no C++ types or exceptions cross it.
 * A context owns one HIP stream and every device buffer of the path.  Host arrays are
 * caller-owned.  A context is single-threaded; distinct contexts may run concurrently.
 * Every call returns NBG_OK (0) or a negative NBG_E_* code; nbg_last_error(ctx) gives
 * the message of the last failure on that context.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference repository root, HasanHelal429/OpenGL-Physics).
 */
#ifndef NBG_H
#define NBG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBG_OK 0
#define NBG_E_INVALID (-1)   /* bad argument (null pointer, n out of range, bad params) */
#define NBG_E_OOM (-2)       /* device allocation failed */
#define NBG_E_CAPACITY (-3)  /* a fixed-capacity pool would overflow; nothing was dropped */
#define NBG_E_HIP (-4)       /* HIP runtime error */
#define NBG_E_STATE (-5)     /* call out of order (e.g. step before upload) */

/* Solver kinds, ngrav::Solver (projects/nbody_core/include/ngrav/Solvers.hpp:11-18). */
#define NBG_SOLVER_DIRECT 0
#define NBG_SOLVER_BARNES_HUT 1

/* Multipole order of accepted cells.  The reference's System always adds the
 * quadrupole in 3D (projects/nbody_core/include/ngrav/System.hpp:59); a tree without
 * ComputeQuadrupoles() is monopole-only (projects/nbody_core/src/SelfTest.cpp:238-240). */
#define NBG_MONOPOLE 0
#define NBG_QUADRUPOLE 1

/* Softening kinds, ngrav::SofteningKind (projects/nbody_core/include/ngrav/Softening.hpp:30-33). */
#define NBG_SOFT_PLUMMER 0
#define NBG_SOFT_SPLINE 1   /* cubic-spline kernel of length h = sqrt(eps2) (Softening::Spline(h),
                               Softening.hpp:132-138), exactly Newtonian beyond 2h */

/* Opening criteria, ngrav::MacKind (projects/nbody_core/include/ngrav/Mac.hpp:13-16). */
#define NBG_MAC_GEOMETRIC 0
#define NBG_MAC_RELATIVE 1  /* per-particle acceleration criterion with the previous |a| */

/* flags */
#define NBG_FLAG_KAHAN 1u           /* compensated FP32 force sums */
#define NBG_FLAG_COUNT 2u           /* record interaction counts: per leaf group for the grouped walk
                                       (nbg_interaction_counts), totals in nbg_stats for both walks */
#define NBG_FLAG_LEARNED 4u         /* add the learned far-field correction (nbg_learn_set_weights) */
#define NBG_FLAG_NO_GRAPH 8u        /* nbg_step: launch eagerly instead of replaying a hipGraph */
#define NBG_FLAG_STAGE_TIMES 16u    /* nbg_step: per-stage HIP-event times summed over all steps
                                       (one host synchronisation per step) */
#define NBG_FLAG_PER_PARTICLE 32u   /* Barnes-Hut with one tree descent per particle
                                       (ComputeAccelBarnesHutPerParticle, Solvers.cpp:359-382)
                                       instead of the grouped walk; implied by NBG_MAC_RELATIVE */

typedef struct nbg_ctx nbg_ctx;

/* Everything one force evaluation and one step need; ngrav::StepParams
 * (projects/nbody_core/include/ngrav/Solvers.hpp:22-40) without the FMM orders.
 * Fields after learn_band were appended later; zero-initialised they select Plummer
 * softening, the geometric MAC and fixed global steps. */
typedef struct nbg_params {
    int solver;     /* NBG_SOLVER_* */
    int multipole;  /* NBG_MONOPOLE | NBG_QUADRUPOLE */
    double G;       /* StepParams::G */
    double eps2;    /* Softening::eps2 (Plummer) */
    double theta;   /* MacParams::theta (geometric MAC) */
    double dt;      /* StepParams::dt (nbg_step only) */
    uint32_t flags; /* NBG_FLAG_* */
    double learn_band; /* NBG_FLAG_LEARNED: an accepted cell gets the learned correction when
                          size / (distance - group radius) exceeds learn_band (0 = every
                          accepted cell).  No reference counterpart. */
    int32_t softening;  /* NBG_SOFT_* (StepParams::soft.kind) */
    int32_t mac;        /* NBG_MAC_* (MacParams::kind) */
    double alpha;       /* MacParams::alpha: relative-MAC force-error target */
    int32_t adaptive;   /* StepParams::adaptive: global dt = eta min sqrt(h/|a|), clamped to
                           [dt/4096, dt] (ChooseAdaptiveDt, Integrator.cpp:41-59) */
    int32_t block;      /* StepParams::block: power-of-2 rung steps (System::BlockStep,
                           System.hpp:239-294); takes precedence over adaptive */
    double eta;         /* StepParams::eta */
    int32_t block_max_rung; /* StepParams::blockMaxRung */
    int32_t reserved;   /* 0 */
    double learn_amin;  /* NBG_FLAG_LEARNED: an accepted cell in the band also needs a correction
                           scale G M s^3 / r^5 >= learn_amin (r: cell to leaf-group centre of
                           mass) to be corrected (0 = no such limit).  No reference counterpart. */
} nbg_params;

/* Opt-in stage timings and tree statistics, like GpuBarnesHutStats
 * (projects/nbody_core/include/ngrav/gpu/GpuBarnesHut.hpp:23-31).  Times are HIP-event
 * milliseconds of the last force evaluation (or of all steps, for nbg_step). */
typedef struct nbg_stats {
    int64_t n_nodes;      /* live tree cells */
    int64_t n_leaves;
    int32_t max_depth;    /* deepest leaf */
    int32_t deep_keys;    /* 1 if the 40-level (two-word) key path was needed */
    int64_t n_p2p;        /* particle-particle interactions (NBG_FLAG_COUNT) */
    int64_t n_m2p;        /* cell-particle interactions (NBG_FLAG_COUNT) */
    double ms_keys;       /* bbox + root + Morton keys */
    double ms_sort;       /* radix sort */
    double ms_build;      /* topology + leaf order + gather */
    double ms_upsweep;    /* moments */
    double ms_walk;       /* tree walk + P2P/M2P (or direct sum) */
    double ms_integrate;  /* kick/drift kernels */
    double ms_total;
    int64_t n_learned;    /* (leaf, cell) pairs given the learned correction (last evaluation) */
    double ms_learned;    /* learned-correction kernels (inside ms_walk's evaluation) */
    double t_advanced;    /* nbg_step: simulated time advanced (sum of the dt System::Step returns) */
    double dt_last;       /* nbg_step: dt of the last step (the coarse dt for block steps) */
    int64_t particle_evals; /* per-particle force evaluations (System::ParticleEvals) */
    int32_t rung_max;     /* nbg_step with block: deepest rung of the last coarse step */
    int32_t key_levels;   /* tree levels the last tree's keys held: 16, 21 or 40 (nbg_tree_export's keys) */
} nbg_stats;

/* Create a context on HIP device `device` able to hold up to max_n particles.
 * Replaces the process-global lazily created GL state of
 * projects/nbody_core/src/gpu/GpuBarnesHut.cpp:452-511.  1 <= max_n <= 2^28 - 1024
 * (268,434,432).  Returns NULL on failure. */
nbg_ctx* nbg_create(int device, int64_t max_n);
void nbg_destroy(nbg_ctx* ctx);
const char* nbg_last_error(const nbg_ctx* ctx);

/* Upload particle state (FP32, original particle order).  vx/vy/vz may be NULL (zero).
 * Replaces the FP64->vec4 pack + glNamedBufferSubData of
 * projects/nbody_core/src/gpu/GpuBarnesHut.cpp:525-531.  Accelerations are cleared. */
int nbg_upload(nbg_ctx* ctx, int64_t n, const float* x, const float* y, const float* z, const float* m,
               const float* vx, const float* vy, const float* vz);

/* Upload a batch of independent systems stored back to back: system s owns particles
 * [offsets[s], offsets[s+1]).  offsets has nsys+1 entries, offsets[0] = 0.  Every
 * system gets its own tree and sees only its own particles.  No reference counterpart:
 * the reference runs one System at a time (projects/nbody_core/include/ngrav/System.hpp). */
int nbg_upload_batch(nbg_ctx* ctx, int32_t nsys, const int64_t* offsets, const float* x, const float* y,
                     const float* z, const float* m, const float* vx, const float* vy, const float* vz);

/* Upload accelerations a(t) (e.g. from a previous run); NULL clears them. */
int nbg_upload_accel(nbg_ctx* ctx, const float* ax, const float* ay, const float* az);

/* Force evaluation a(x) into the context's acceleration buffer, G included.
 * Replaces ngrav::gpu::ComputeAccelGpuBarnesHut (projects/nbody_core/src/gpu/GpuBarnesHut.cpp:515-694)
 * and ngrav::gpu::ComputeAccelGpuDirect (projects/nbody_core/src/gpu/GpuDirect.cpp:61-117), and on the
 * CPU side System::ComputeAccel (projects/nbody_core/include/ngrav/System.hpp:43-78), whose
 * Barnes-Hut case is the grouped walk ComputeAccelBarnesHutGroup
 * (projects/nbody_core/src/Solvers.cpp:279-355) over AdaptiveTree<3> (projects/nbody_core/src/Tree.cpp).
 * Synchronous: returns when the accelerations are in device memory. */
int nbg_accel(nbg_ctx* ctx, const nbg_params* p, nbg_stats* stats);

/* nsteps kick-drift-kick leapfrog steps of p->dt with force evaluations by p->solver.
 * LeapfrogStep (projects/nbody_core/src/Integrator.cpp:9-39) as driven by System::Step
 * (projects/nbody_core/include/ngrav/System.hpp:89-108).  Precondition: the context holds
 * a(t) (call nbg_accel first, as System::Prime does).  Unless NBG_FLAG_NO_GRAPH is set,
 * one step is captured once into a hipGraph and replayed.
 *   p->adaptive: each step's dt is chosen on the device from a(t) (ChooseAdaptiveDt,
 *     Integrator.cpp:41-59) and read by the kick/drift kernels; graph-captured too.
 *   p->block: each call of System::BlockStep (System.hpp:239-294) is one coarse step of
 *     p->dt split into power-of-2 rungs; every substep rebuilds the tree and evaluates
 *     the closing set with the per-particle geometric walk (Barnes-Hut) or the direct
 *     sum onto those targets.  Eager (the substep count depends on the rungs).
 * stats->t_advanced / dt_last / particle_evals / rung_max report what the steps took. */
int nbg_step(nbg_ctx* ctx, const nbg_params* p, int32_t nsteps, nbg_stats* stats);

/* Copy state back to the host (original particle order).  Any pointer may be NULL.
 * Replaces the readback + FP32->FP64 unpack of projects/nbody_core/src/gpu/GpuBarnesHut.cpp:672-679. */
int nbg_download(nbg_ctx* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* ax,
                 float* ay, float* az);

/* Tree of the last Barnes-Hut evaluation, for bit-exact comparison with AdaptiveTree<3>
 * (projects/nbody_core/include/ngrav/Tree.hpp:44-63).  Nodes are numbered in DFS pre-order
 * with children in ascending octant, the order AdaptiveTree::Finalize visits them
 * (projects/nbody_core/src/Tree.cpp:157-175).  Any output pointer may be NULL.
 *   n_nodes  in: capacity of the node arrays; out: node count
 *   keys     n uint64, the sorted 63-bit keys (levels 1..21, 3 bits per level, MSB first)
 *   order    n int32, AdaptiveTree::order (leaf members in ascending original index)
 *   center   4 doubles per node: centre x,y,z and halfSize (FP64, bit-exact)
 *   topo     5 int32 per node: parent, firstParticle (-1 internal), particleCount
 *            (0 internal), depth, childCount
 *   children 8 int32 per node (-1 = empty octant)
 *   moments  10 floats per node: com x,y,z, mass, maxRadius, Qxx,Qyy,Qxy,Qxz,Qyz */
int nbg_tree_export(nbg_ctx* ctx, int64_t* n_nodes, uint64_t* keys, int32_t* order, double* center,
                    int32_t* topo, int32_t* children, float* moments);

/* Per-leaf interaction counts of the last Barnes-Hut evaluation run with
 * NBG_FLAG_COUNT, indexed like nbg_tree_export (zero for internal nodes): far cells
 * (M2P sources) and near particles (P2P sources, self excluded) of each leaf group,
 * as CollectGroupInteractions builds them (projects/nbody_core/src/Solvers.cpp:248-277). */
int nbg_interaction_counts(nbg_ctx* ctx, int64_t n_nodes, int64_t* n_far, int64_t* n_near);

/* FP64 diagnostics of the current state: out[0]=kinetic, out[1]=potential (O(N^2), Plummer
 * softened), out[2..4]=centre of mass, out[5..7]=angular momentum about the COM,
 * out[8..10]=linear momentum.  System::TotalEnergy / CenterOfMass / AngularMomentum
 * (projects/nbody_core/include/ngrav/System.hpp:119-193) and the momentum of
 * DeckSim::Snapshot (projects/nbody_core/src/DeckSim.cpp:147-153). For a batch, system 0.
 * potential is skipped (left 0) when with_potential is 0. */
int nbg_diagnostics(nbg_ctx* ctx, const nbg_params* p, int32_t with_potential, double* out);

/* Direct-sum accelerations (G included) onto a subset of targets, all particles of the
 * target's system as sources; FP32 with compensated sums.  ComputeAccelDirectTargets
 * (projects/nbody_core/src/Solvers.cpp:108-134).  Results go to host arrays of length nt. */
int nbg_accel_direct_targets(nbg_ctx* ctx, const nbg_params* p, int64_t nt, const int64_t* targets, float* ax,
                             float* ay, float* az);

/* Barnes-Hut accelerations onto a subset of targets by the per-particle geometric walk,
 * all particles as sources, after a fresh tree build: ComputeAccelBarnesHutTargets
 * (projects/nbody_core/src/Solvers.cpp:384-405), the force pass of block steps.
 * Results go to host arrays of length nt; the context's accelerations are unchanged. */
int nbg_accel_bh_targets(nbg_ctx* ctx, const nbg_params* p, int64_t nt, const int64_t* targets, float* ax,
                         float* ay, float* az);

/* Per-particle rungs of the last block step (System::Rungs, System.hpp:203-205), original
 * order; zeros before any block step. */
int nbg_rungs(nbg_ctx* ctx, int32_t* rungs);

/* Launch choices of the force kernels, set directly or by nbg_autotune.  No reference
 * counterpart: the reference's GL kernels have fixed work-group sizes
 * (projects/nbody_core/src/gpu/GpuBarnesHut.cpp, GpuDirect.cpp). */
typedef struct nbg_tuning {
    int32_t walk_kernel; /* 0: bin walk (one wavefront per 64 sorted targets), 1: one wavefront
                            per leaf group (no learned-pair listing; learned runs use the bin walk) */
    int32_t near_tile;   /* bin walk: near-field P2P sources loaded per batch: 4, 8 or 16 (default
                            16; the learned walk takes 16 or, for 4 and 8, 8) */
    int32_t direct_tile; /* direct sum: sources loaded per batch: 2, 4 or 8 */
    int32_t reserved;    /* 0 */
} nbg_tuning;

int nbg_tuning_get(nbg_ctx* ctx, nbg_tuning* t);
/* Drops the captured step graph (it holds the previous kernels). */
int nbg_tuning_set(nbg_ctx* ctx, const nbg_tuning* t);

/* Time every candidate tuning of p->solver's force kernel on the uploaded state (one
 * untimed and `reps` timed evaluations each, HIP-event kernel time), keep the fastest in
 * the context and return it in *best.  ms (nullable, cap entries) receives the mean time
 * of each candidate in candidate order, *n_out the number of candidates.  Candidates: bin
 * walk with near_tile 4/8/16 and, for Plummer softening with the grouped geometric walk
 * and no learned correction, the leaf-group walk (Barnes-Hut); direct_tile 2/4/8
 * (direct).  The accelerations are overwritten with those of the last evaluation. */
int nbg_autotune(nbg_ctx* ctx, const nbg_params* p, int32_t reps, nbg_tuning* best, double* ms, int32_t cap,
                 int32_t* n_out);

/* nbg_diagnostics for system `sys` of a batch (nbg_upload_batch). */
int nbg_diagnostics_system(nbg_ctx* ctx, const nbg_params* p, int32_t sys, int32_t with_potential, double* out);

/* ---------------------------------------------------------------------------------
 * Learned far-field correction.  No reference counterpart: the reference has no learned
 * component.  A multilayer perceptron (13 inputs, three hidden layers of 64 with SiLU or
 * ReLU, 3 outputs) maps the geometry and moments of one accepted (leaf group, cell) pair
 * to the residual of that cell's multipole force against direct summation, averaged over
 * the group's members; the walk adds the prediction to every member of the group.
 * Inference and training run in BF16 on the matrix cores with FP32 accumulation.
 *
 * Model blob (NBG_LEARN_NPARAM floats): feature mean[16], feature 1/std[16], then
 * W1[13][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][3] b4[3] (y = x W + b).
 * Features 13..15 do not exist; the unused mean[NBG_LEARN_ACT_SLOT] holds the model's
 * activation (NBG_LEARN_ACT_SILU = 0, NBG_LEARN_ACT_RELU = 1), which nbg_learn_standardize
 * keeps.
 * ------------------------------------------------------------------------------- */
#define NBG_LEARN_FEATURES 13
#define NBG_LEARN_HIDDEN 64
#define NBG_LEARN_NTRAIN 9411 /* trainable parameters */
#define NBG_LEARN_NPARAM (32 + NBG_LEARN_NTRAIN)
#define NBG_LEARN_ACT_SLOT 15
#define NBG_LEARN_ACT_SILU 0
#define NBG_LEARN_ACT_RELU 1

/* Install a model blob (FP32 master copy; BF16 copy derived).  Required before
 * NBG_FLAG_LEARNED evaluations and before nbg_learn_train.  NBG_E_INVALID unless
 * blob[NBG_LEARN_ACT_SLOT] is one of the NBG_LEARN_ACT_* codes. */
int nbg_learn_set_weights(nbg_ctx* ctx, const float* blob, int64_t n);
/* Copy the current model blob (after training) to the host. */
int nbg_learn_get_weights(nbg_ctx* ctx, float* blob, int64_t n);

/* Append training samples from the uploaded snapshot to the context's sample pool:
 * direct-sum accelerations, then a Barnes-Hut walk with p->theta and p->learn_band that
 * lists its accepted (leaf, cell) pairs, then for each pair (up to max_new) its 13
 * features, the exact residual (FP64 direct sum over the cell minus the walk's multipole
 * force, averaged over the leaf's members, divided by the pair's scale) and a loss weight.
 * *n_pool receives the pool size.  The context's positions, velocities and accelerations
 * are left as they were (the evaluations it makes are scratch), so a primed nbg_step may
 * follow directly. */
int nbg_learn_collect(nbg_ctx* ctx, const nbg_params* p, int64_t max_new, int64_t* n_pool);
int nbg_learn_clear(nbg_ctx* ctx);
/* Download pool samples [offset, offset+count): feats 16 floats per sample (0..12 the
 * features, 13 and 14 the leaf and cell node ids as int32 bits, numbered as
 * nbg_tree_export numbers them), targets 4 floats (residual x,y,z and loss weight). */
int nbg_learn_pool(nbg_ctx* ctx, int64_t offset, int64_t count, float* feats, float* targets);
/* Set the feature standardisation of the installed model from the pool's mean and
 * standard deviation (first `count` samples, all if <= 0). */
int nbg_learn_standardize(nbg_ctx* ctx, int64_t count);
/* Adam on the pool: `steps` steps of `batch` samples (multiple of 64) drawn with
 * replacement by a counter-based hash of `seed`.  loss (nullable, `steps` floats)
 * receives the weighted mean-square loss of each step's batch before its update. */
int nbg_learn_train(nbg_ctx* ctx, int32_t steps, int32_t batch, float lr, uint64_t seed, float* loss);
/* Forward + backward of the installed model on n caller samples (feats 16 floats,
 * targets 4 floats each), through the same kernels as training, without an update:
 * out (n x 3) predictions, *loss = sum_i w_i |y_i - t_i|^2 / n, grad (NBG_LEARN_NTRAIN)
 * d loss / d parameters in blob order (without the standardisation entries). */
int nbg_learn_grad(nbg_ctx* ctx, int64_t n, const float* feats, const float* targets, float* out, float* loss,
                   float* grad);

/* Upload with FP64 positions; masses and velocities FP32.  No counterpart in the
 * reference's GPU path; the reference's own state is FP64 (State.hpp:16-109) and this is
 * the way to give the device that state without rounding positions to FP32.  The tree is
 * built from the FP64 positions (root box, keys and centres of mass in FP64, as
 * Tree.cpp:33-254 does), and the grouped walk forms every pair difference in FP32 from
 * positions relative to FP64 points near the pair (sources relative to their leaf's
 * centre, targets and cells' centres of mass relative to the centre of the walk bin's
 * first leaf, plus the rounded offsets between those points), so no absolute position is
 * rounded to FP32 on the way to a force.  The direct sum, the per-particle walks and the
 * learned correction take the FP32 roundings of the positions.  Force evaluations only:
 * nbg_step returns NBG_E_STATE.  nbg_upload / nbg_upload_batch return to an FP32 state. */
int nbg_upload_f64(nbg_ctx* ctx, int64_t n, const double* x, const double* y, const double* z, const float* m,
                   const float* vx, const float* vy, const float* vz);

/* No counterpart in the reference.  For callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * The direct sum of nbg_accel with NBG_SOLVER_DIRECT (same kernels, same results) on
 * caller-owned DEVICE arrays of one system: x, y, z, m in, ax, ay, az out (G included),
 * enqueued on `stream` (a hipStream_t; NULL = the default stream) of HIP device `device`,
 * which the call makes current; it returns once the work is enqueued.  Uses p->G,
 * p->eps2, p->softening and NBG_FLAG_KAHAN; p->solver must be NBG_SOLVER_DIRECT.
 * No context; calls no hipMalloc, hipFree or allocating routine: the packed source
 * records are staged in a library-wide array of NBG_NOALLOC_MAX_N records (16 MiB) that
 * each device holds from its first call until the process ends, so
 * 1 <= n <= NBG_NOALLOC_MAX_N (NBG_E_INVALID otherwise).  Calls whose work overlaps on
 * one device (two host threads, or two streams) race on that array.
 * Returns NBG_OK, NBG_E_INVALID (null pointer, n out of range, bad params or device) or
 * NBG_E_HIP (the symbol lookup or the launch failed). */
#define NBG_NOALLOC_MAX_N (1 << 20)
int nbg_accel_direct_noalloc(int device, void* stream, const nbg_params* p, int64_t n, const float* x,
                             const float* y, const float* z, const float* m, float* ax, float* ay, float* az);

#ifdef __cplusplus
}
#endif

#endif /* NBG_H */
