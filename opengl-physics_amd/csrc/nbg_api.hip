// nbg_api.hip -- the C ABI of libnbg_hip.so (declared in include/nbg.h): context and
// device memory, the force-evaluation pipeline, the leapfrog driver and its hipGraph.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "nbg_internal.hpp"

namespace nbg {
void launch_pack4(nbg_ctx* c);  // input-order records for the per-leaf gather
}

using namespace nbg;

namespace {

constexpr int kMaxSys = 1 << 20;
constexpr int kMaxRung = 24;  // 2^24 substeps per coarse step at most

template <typename T>
int alloc_into(nbg_ctx* c, T*& p, size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    hipError_t e;
    p = dev_alloc<T>(count, e);
    if (!p) {
        c->set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return NBG_E_OOM;
    }
    return NBG_OK;
}

// One slot (0: com, 1: quad, 3: aux) of the first n node records, to host memory.
hipError_t copy_rec_slot(const nbg_ctx* c, float4* dst, int slot, int64_t n) {
    if (n <= 0) return hipSuccess;
    return hipMemcpy2D(dst, sizeof(float4), c->nd_rec + slot, 4 * sizeof(float4), sizeof(float4),
                       static_cast<size_t>(n), hipMemcpyDeviceToHost);
}

template <typename T>
void free_dev(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int check_launch(nbg_ctx* c) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        c->set_error(std::string("kernel launch: ") + hipGetErrorString(e));
        return NBG_E_HIP;
    }
    return NBG_OK;
}

int validate(nbg_ctx* c, const nbg_params* p) {
    if (!p) {
        c->set_error("params is NULL");
        return NBG_E_INVALID;
    }
    if (p->solver != NBG_SOLVER_DIRECT && p->solver != NBG_SOLVER_BARNES_HUT) {
        c->set_error("unknown solver");
        return NBG_E_INVALID;
    }
    if (p->multipole != NBG_MONOPOLE && p->multipole != NBG_QUADRUPOLE) {
        c->set_error("unknown multipole order");
        return NBG_E_INVALID;
    }
    if (!(p->eps2 >= 0.0) || !(p->theta >= 0.0) || !(p->G == p->G)) {
        c->set_error("eps2 and theta must be >= 0");
        return NBG_E_INVALID;
    }
    if (!c->have_state) {
        c->set_error("no particles uploaded");
        return NBG_E_STATE;
    }
    if (p->softening != NBG_SOFT_PLUMMER && p->softening != NBG_SOFT_SPLINE) {
        c->set_error("unknown softening kind");
        return NBG_E_INVALID;
    }
    if (p->softening == NBG_SOFT_SPLINE && !(p->eps2 > 0.0)) {
        c->set_error("spline softening needs a softening length h = sqrt(eps2) > 0");
        return NBG_E_INVALID;
    }
    if (p->mac != NBG_MAC_GEOMETRIC && p->mac != NBG_MAC_RELATIVE) {
        c->set_error("unknown opening criterion (mac)");
        return NBG_E_INVALID;
    }
    if (p->mac == NBG_MAC_RELATIVE && !(p->alpha >= 0.0 && p->alpha < 1e300)) {
        c->set_error("the relative MAC needs a finite alpha >= 0");
        return NBG_E_INVALID;
    }
    if ((p->adaptive || p->block) && !(p->eta > 0.0 && p->dt > 0.0)) {
        c->set_error("adaptive and block steps need eta > 0 and dt > 0");
        return NBG_E_INVALID;
    }
    if (p->block && p->block_max_rung > kMaxRung) {
        c->set_error("block_max_rung must be <= " + std::to_string(kMaxRung));
        return NBG_E_INVALID;
    }
    if (p->flags & NBG_FLAG_LEARNED) {
        if (p->softening != NBG_SOFT_PLUMMER || p->mac != NBG_MAC_GEOMETRIC || (p->flags & NBG_FLAG_PER_PARTICLE) ||
            p->block) {
            c->set_error("the learned correction runs with Plummer softening, the grouped geometric walk and global steps");
            return NBG_E_INVALID;
        }
        if (p->solver != NBG_SOLVER_BARNES_HUT || p->multipole != NBG_QUADRUPOLE) {
            c->set_error("the learned correction needs the Barnes-Hut solver with quadrupoles");
            return NBG_E_INVALID;
        }
        if (!c->have_weights) {
            c->set_error("no learned model installed (nbg_learn_set_weights)");
            return NBG_E_STATE;
        }
        if (!(p->learn_band >= 0.0) || !(p->learn_amin >= 0.0)) {
            c->set_error("learn_band and learn_amin must be >= 0");
            return NBG_E_INVALID;
        }
        if (int r = learn_alloc(c)) return r;
    }
    return NBG_OK;
}

// Record a stage event; while capturing a graph it becomes an event-record node.
inline void rec(nbg_ctx* c, int i) {
    if (c->capturing)
        (void)hipEventRecordWithFlags(c->ev[i], c->stream, hipEventRecordExternal);
    else
        (void)hipEventRecord(c->ev[i], c->stream);
}

void drop_graph(nbg_ctx* c);

// Enqueue one Barnes-Hut force evaluation (no host synchronisation).
// The key levels to retry with after a tree at `levels` overflowed (a cell at the deepest
// level held more than ncrit particles).  An overflow at 13 or 16 levels raises the
// context's first try to the next tier (16, 21) until the next upload; the two-word keys
// are taken per evaluation.
int next_levels(nbg_ctx* c, int levels) {
    if (levels == kLevelsTiny) {
        c->key_levels = kLevelsFast;
        return kLevelsFast;
    }
    if (levels == kLevelsFast) {
        c->key_levels = kLevelsShort;
        return kLevelsShort;
    }
    return kLevelsDeep;
}

void enqueue_bh(nbg_ctx* c, const nbg_params& p, int levels, bool timed, bool list_only = false) {
    const bool learn = (p.flags & NBG_FLAG_LEARNED) != 0;
    if (timed) rec(c, 0);
    launch_bbox_and_keys(c, levels);
    if (timed) rec(c, 1);
    launch_sort(c, levels);
    if (timed) rec(c, 2);
    launch_build(c, levels);
    launch_pack4(c);
    if (timed) rec(c, 3);
    launch_upsweep(c, levels, p.multipole, p.G, learn);
    if (c->f64_state) launch_rel(c);
    if (timed) rec(c, 4);
    if ((p.flags & NBG_FLAG_COUNT) && c->cnt_far) {
        launch_fill_i32(c, c->cnt_far, c->cap_nodes, 0);
        launch_fill_i32(c, c->cnt_near, c->cap_nodes, 0);
    }
    launch_walk(c, p);
    if (timed) rec(c, 5);
    if (learn && !list_only) {
        if (timed) rec(c, 12);
        launch_learn_zero(c);
        launch_learn_correct(c, p);
        if (timed) rec(c, 13);
    } else if (learn) {
        // the listing walk's own accelerations, uncorrected, to input order
        launch_learn_zero(c);
        launch_learn_apply(c);
    }
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

int read_scalars(nbg_ctx* c) {
    NBG_HIP_CHECK(c, hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return NBG_OK;
}

void fill_tree_stats(nbg_ctx* c, nbg_stats* s) {
    s->n_nodes = c->hsc->n_nodes;
    s->n_leaves = c->hsc->n_leaves;
    s->max_depth = c->hsc->max_depth;
    s->deep_keys = (c->last_levels == kLevelsDeep) ? 1 : 0;
    s->key_levels = c->last_levels;
}

// One full force evaluation with the short-key path, falling back to two-word keys
// when a depth-21 cell still holds more than ncrit particles.
int accel_eager(nbg_ctx* c, const nbg_params& p, nbg_stats* stats, bool list_only = false) {
    const bool timed = stats != nullptr;
    const bool learn = (p.flags & NBG_FLAG_LEARNED) != 0;
    c->last_G = p.G;
    if (p.solver == NBG_SOLVER_DIRECT) {
        if (timed) rec(c, 0);
        launch_direct(c, p);
        if (timed) rec(c, 5);
        if (int r = check_launch(c)) return r;
        NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        c->have_tree = false;
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->ms_walk = elapsed(c->ev[0], c->ev[5]);
            stats->ms_total = stats->ms_walk;
        }
        return NBG_OK;
    }
    // the relative MAC reads |a| of the previous evaluation: capture it once, before the
    // walk (a deep-key re-run below must not see this evaluation's accelerations)
    if (p.mac == NBG_MAC_RELATIVE) launch_aold(c);
    int levels = c->key_levels;
    for (;;) {
        enqueue_bh(c, p, levels, timed, list_only);
        if (int r = check_launch(c)) return r;
        if (int r = read_scalars(c)) return r;
        if (c->hsc->cap_exceeded) {
            c->set_error("node pool capacity exceeded");
            return NBG_E_CAPACITY;
        }
        if (c->hsc->overflow && levels < kLevelsDeep) {
            // a cell at the keys' deepest level holds more than ncrit particles: more levels
            // (13 -> 16 -> 21, kept for this state; 21 -> the two-word keys, for this evaluation)
            levels = next_levels(c, levels);
            continue;
        }
        if (learn && static_cast<int64_t>(c->hsc->learn_pairs) > c->lpairs_cap) {
            // the pair list was too short: grow it and evaluate again (the state is unchanged)
            drop_graph(c);
            if (int r = learn_grow_pairs(c, c->hsc->learn_pairs)) return r;
            continue;
        }
        if (learn && !list_only && c->hsc->learn_fix_over && !c->learn_two_pass) {
            // a correction outside the one-pass fixed-point range: this context sums them
            // with per-leaf units from now on (graphs are captured again)
            drop_graph(c);
            c->learn_two_pass = true;
            continue;
        }
        break;
    }
    c->last_levels = levels;
    c->have_tree = true;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->particle_evals = c->n;
        fill_tree_stats(c, stats);
        stats->ms_keys = elapsed(c->ev[0], c->ev[1]);
        stats->ms_sort = elapsed(c->ev[1], c->ev[2]);
        stats->ms_build = elapsed(c->ev[2], c->ev[3]);
        stats->ms_upsweep = elapsed(c->ev[3], c->ev[4]);
        stats->ms_walk = elapsed(c->ev[4], c->ev[5]);
        stats->ms_total = elapsed(c->ev[0], c->ev[5]);
        if (learn) {
            stats->n_learned = c->hsc->learn_pairs;
            if (!list_only) {
                stats->ms_learned = elapsed(c->ev[12], c->ev[13]);
                stats->ms_total = elapsed(c->ev[0], c->ev[13]);
            }
        }
        if ((p.flags & NBG_FLAG_COUNT) && (p.mac == NBG_MAC_RELATIVE || (p.flags & NBG_FLAG_PER_PARTICLE))) {
            // the per-particle walk counts its interactions directly
            stats->n_m2p = static_cast<int64_t>(c->hsc->pp_m2p);
            stats->n_p2p = static_cast<int64_t>(c->hsc->pp_p2p);
        } else if (p.flags & NBG_FLAG_COUNT) {
            // interaction totals: per group counts times group sizes
            const int nn = c->hsc->n_nodes;
            std::vector<int32_t> f(nn), q(nn), cnt(nn);
            std::vector<float4> aux(nn);
            (void)hipMemcpy(f.data(), c->cnt_far, nn * sizeof(int32_t), hipMemcpyDeviceToHost);
            (void)hipMemcpy(q.data(), c->cnt_near, nn * sizeof(int32_t), hipMemcpyDeviceToHost);
            (void)hipMemcpy(cnt.data(), c->nd_count, nn * sizeof(int32_t), hipMemcpyDeviceToHost);
            (void)copy_rec_slot(c, aux.data(), 3, nn);
            int64_t m2p = 0, p2p = 0;
            for (int i = 0; i < nn; ++i) {
                uint32_t info;
                std::memcpy(&info, &aux[i].w, 4);
                if (!(info & kLeafBit)) continue;
                m2p += static_cast<int64_t>(f[i]) * cnt[i];
                p2p += static_cast<int64_t>(q[i]) * cnt[i];
            }
            stats->n_m2p = m2p;
            stats->n_p2p = p2p;
        }
    }
    return NBG_OK;
}

// One KDK step, enqueued (force evaluation with `levels`).  With `timed`, event
// records bracket every stage (ev[8..9] first kick + drift, ev[0..5] force stages,
// ev[10..11] closing kick).
void enqueue_step(nbg_ctx* c, const nbg_params& p, int levels, bool timed) {
    const float dt = static_cast<float>(p.dt);
    const float hdt = static_cast<float>(0.5 * p.dt);
    if (timed) rec(c, 8);
    if (p.adaptive) {
        launch_adaptive_dt(c, p);
        launch_kick_drift_dev(c);
    } else {
        launch_kick_drift(c, hdt, dt);
    }
    if (timed) rec(c, 9);
    if (p.solver == NBG_SOLVER_DIRECT) {
        if (timed) rec(c, 4);
        launch_direct(c, p);
        if (timed) rec(c, 5);
    } else {
        nbg_params q = p;
        q.flags &= ~NBG_FLAG_COUNT;
        if (p.mac == NBG_MAC_RELATIVE) launch_aold(c);
        enqueue_bh(c, q, levels, timed);
    }
    if (timed) rec(c, 10);
    if (p.adaptive)
        launch_kick_dev(c);
    else
        launch_kick(c, hdt);
    if (timed) rec(c, 11);
}

void add_stage_times(nbg_ctx* c, const nbg_params& p, nbg_stats* s) {
    s->ms_integrate += elapsed(c->ev[8], c->ev[9]) + elapsed(c->ev[10], c->ev[11]);
    s->ms_walk += elapsed(c->ev[4], c->ev[5]);
    if (p.solver == NBG_SOLVER_BARNES_HUT) {
        s->ms_keys += elapsed(c->ev[0], c->ev[1]);
        s->ms_sort += elapsed(c->ev[1], c->ev[2]);
        s->ms_build += elapsed(c->ev[2], c->ev[3]);
        s->ms_upsweep += elapsed(c->ev[3], c->ev[4]);
        if (p.flags & NBG_FLAG_LEARNED) s->ms_learned += elapsed(c->ev[12], c->ev[13]);
    }
}

int save_backup(nbg_ctx* c) {
    float* src[9] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az};
    for (int k = 0; k < 9; ++k) launch_copy_f32(c, c->backup + static_cast<int64_t>(k) * c->cap_n, src[k], c->n);
    return check_launch(c);
}

int restore_backup(nbg_ctx* c) {
    float* dst[9] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az};
    for (int k = 0; k < 9; ++k) launch_copy_f32(c, dst[k], c->backup + static_cast<int64_t>(k) * c->cap_n, c->n);
    return check_launch(c);
}

void drop_graph(nbg_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_valid = false;
}

bool same_params(const nbg_params& a, const nbg_params& b) {
    return a.solver == b.solver && a.multipole == b.multipole && a.G == b.G && a.eps2 == b.eps2 &&
           a.theta == b.theta && a.dt == b.dt && (a.flags & ~NBG_FLAG_COUNT) == (b.flags & ~NBG_FLAG_COUNT) &&
           a.learn_band == b.learn_band && a.softening == b.softening && a.mac == b.mac && a.alpha == b.alpha &&
           a.adaptive == b.adaptive && a.eta == b.eta && a.learn_amin == b.learn_amin;
}

// Eager steps with a scalar read-back per step to catch the deep-key case.
int steps_eager(nbg_ctx* c, const nbg_params& p, int nsteps, nbg_stats* stats) {
    const float dt = static_cast<float>(p.dt);
    const float hdt = static_cast<float>(0.5 * p.dt);
    const bool timed = (p.flags & NBG_FLAG_STAGE_TIMES) && stats;
    for (int s = 0; s < nsteps; ++s) {
        if (timed) rec(c, 8);
        if (p.adaptive) {
            launch_adaptive_dt(c, p);
            launch_kick_drift_dev(c);
        } else {
            launch_kick_drift(c, hdt, dt);
        }
        if (timed) rec(c, 9);
        if (int r = check_launch(c)) return r;
        nbg_params q = p;
        q.flags &= ~NBG_FLAG_COUNT;
        nbg_stats one;
        if (int r = accel_eager(c, q, timed ? &one : nullptr)) return r;
        if (timed) rec(c, 10);
        if (p.adaptive)
            launch_kick_dev(c);
        else
            launch_kick(c, hdt);
        if (timed) rec(c, 11);
        if (int r = check_launch(c)) return r;
        if (timed) {
            NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
            stats->ms_keys += one.ms_keys;
            stats->ms_sort += one.ms_sort;
            stats->ms_build += one.ms_build;
            stats->ms_upsweep += one.ms_upsweep;
            stats->ms_walk += one.ms_walk;
            stats->ms_learned += one.ms_learned;
            stats->ms_integrate += elapsed(c->ev[8], c->ev[9]) + elapsed(c->ev[10], c->ev[11]);
        }
    }
    return NBG_OK;
}

// Tree of the current positions (keys, sort, build, gather, moments), with the two-word
// key re-run when a depth-21 cell still holds more than ncrit particles.
int build_tree_eager(nbg_ctx* c, const nbg_params& p) {
    int levels = c->key_levels;
    for (;;) {
        launch_bbox_and_keys(c, levels);
        launch_sort(c, levels);
        launch_build(c, levels);
        launch_pack4(c);
        launch_upsweep(c, levels, p.multipole, p.G, false);
        if (int r = check_launch(c)) return r;
        if (int r = read_scalars(c)) return r;
        if (c->hsc->cap_exceeded) {
            c->set_error("node pool capacity exceeded");
            return NBG_E_CAPACITY;
        }
        if (c->hsc->overflow && levels < kLevelsDeep) {
            levels = next_levels(c, levels);
            continue;
        }
        break;
    }
    c->last_levels = levels;
    c->last_G = p.G;
    c->have_tree = true;
    return NBG_OK;
}

// nsteps calls of System::BlockStep (System.hpp:239-294): rungs from a(t), then 2^rmax
// substeps of opening kicks, a global drift, a force pass over the closing set (fresh
// tree + per-particle geometric walk, or the direct sum onto those targets) and closing
// kicks.  The last substep's closing set is every particle, so a(t + dt) is complete.
int block_steps(nbg_ctx* c, const nbg_params& p, int nsteps, nbg_stats* stats) {
    const double base = p.dt;
    const int max_rung = std::max(0, p.block_max_rung);
    nbg_params q = p;
    q.flags &= ~(NBG_FLAG_COUNT | NBG_FLAG_LEARNED);
    q.mac = NBG_MAC_GEOMETRIC;  // ComputeAccelBarnesHutTargets always takes the geometric MAC
    q.block_max_rung = max_rung;
    NBG_HIP_CHECK(c, hipMemsetAsync(&c->dsc->particle_evals, 0, sizeof(uint64_t), c->stream));
    int rmax = 0;
    for (int s = 0; s < nsteps; ++s) {
        launch_rungs(c, q);
        if (int r = check_launch(c)) return r;
        NBG_HIP_CHECK(c, hipMemcpyAsync(&c->hsc->rung_max, &c->dsc->rung_max, sizeof(int32_t), hipMemcpyDeviceToHost,
                                        c->stream));
        NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        rmax = c->hsc->rung_max;
        c->have_rungs = true;
        const int nsub = 1 << rmax;
        const float dt_min = static_cast<float>(base / nsub);
        for (int sub = 0; sub < nsub; ++sub) {
            launch_block_kick(c, q, sub, rmax, false);
            launch_drift(c, dt_min);
            if (p.solver == NBG_SOLVER_BARNES_HUT) {
                if (int r = build_tree_eager(c, q)) return r;
                launch_active_list(c, sub, rmax, true);
                launch_walk_pp(c, q, c->tlist, &c->dsc->tcount, nullptr);
            } else {
                launch_active_list(c, sub, rmax, false);
                launch_direct_targets(c, q, c->n, c->tlist, &c->dsc->tcount, nullptr);
                c->have_tree = false;
            }
            launch_block_kick(c, q, sub, rmax, true);
            if (int r = check_launch(c)) return r;
        }
    }
    if (stats) {
        NBG_HIP_CHECK(c, hipMemcpyAsync(&c->hsc->particle_evals, &c->dsc->particle_evals, sizeof(uint64_t),
                                        hipMemcpyDeviceToHost, c->stream));
        NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        stats->t_advanced = base * nsteps;
        stats->dt_last = base;
        stats->rung_max = rmax;
        stats->particle_evals = static_cast<int64_t>(c->hsc->particle_evals);
    }
    return NBG_OK;
}

}  // namespace

extern "C" {

nbg_ctx* nbg_create(int device, int64_t max_n) {
    // The force kernels address sorted particles through 32-bit byte offsets (index << 4).
    if (max_n < 1 || max_n > kMaxParticles) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    nbg_ctx* c = new (std::nothrow) nbg_ctx();
    if (!c) return nullptr;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    c->cap_n = max_n;
    if (const char* w = std::getenv("NBG_WALK_KERNEL")) c->tune.walk_kernel = (std::string(w) == "leaf") ? 1 : 0;
    c->cap_nodes = max_n + (40 * max_n + 8) / 9 + 64;
    const size_t N = static_cast<size_t>(max_n);
    const size_t C = static_cast<size_t>(c->cap_nodes);
    int r = NBG_OK;
    r |= alloc_into(c, c->x, N);
    r |= alloc_into(c, c->y, N);
    r |= alloc_into(c, c->z, N);
    r |= alloc_into(c, c->m, N);
    r |= alloc_into(c, c->vx, N);
    r |= alloc_into(c, c->vy, N);
    r |= alloc_into(c, c->vz, N);
    r |= alloc_into(c, c->ax, N);
    r |= alloc_into(c, c->ay, N);
    r |= alloc_into(c, c->az, N);
    r |= alloc_into(c, c->backup, 9 * N);
    r |= alloc_into(c, c->pos_sys, N);
    r |= alloc_into(c, c->key_a, N);
    r |= alloc_into(c, c->key_b, N);
    r |= alloc_into(c, c->klo_a, N);
    r |= alloc_into(c, c->klo_b, N);
    r |= alloc_into(c, c->val_a, N);
    r |= alloc_into(c, c->val_b, N);
    r |= alloc_into(c, c->leaf_depth, N);
    r |= alloc_into(c, c->cpl_prev, N);
    r |= alloc_into(c, c->node_scan, N + 1);
    r |= alloc_into(c, c->blk_depth, N / 256 + 1);
    r |= alloc_into(c, c->leaf_list, N);
    r |= alloc_into(c, c->leaf_of, N);
    r |= alloc_into(c, c->spos, N + kSposPad);
    r |= alloc_into(c, c->stage4, N);
    r |= alloc_into(c, c->nd_start, C);
    r |= alloc_into(c, c->nd_count, C);
    r |= alloc_into(c, c->nd_depth, C);
    r |= alloc_into(c, c->nd_parent, C);
    r |= alloc_into(c, c->nd_oct, C);
    r |= alloc_into(c, c->nd_child, 8 * C);
    r |= alloc_into(c, c->nd_geom, 4 * C);
    r |= alloc_into(c, c->nd_rec, 4 * C);
    r |= alloc_into(c, c->nd_com64, C);
    r |= alloc_into(c, c->nd_rmax64, C);
    r |= alloc_into(c, c->depth_nodes, C);
    r |= alloc_into(c, c->dep_hist, static_cast<size_t>(kMaxDepth) * kNodeBlocks);
    r |= alloc_into(c, c->dep_off, static_cast<size_t>(kMaxDepth) * kNodeBlocks);
    r |= alloc_into(c, c->dep_part, static_cast<size_t>(kMaxDepth) + 8);
    r |= alloc_into(c, c->cnt_far, C);
    r |= alloc_into(c, c->cnt_near, C);
    r |= alloc_into(c, c->aold, N);
    r |= alloc_into(c, c->rung, N);
    r |= alloc_into(c, c->tlist, N);
    r |= alloc_into(c, c->list_off, N / 1024 + 2);
    r |= alloc_into(c, c->red_part, kReduceBlocks);
    r |= alloc_into(c, c->dsc, 1);
    c->diag_part_cap = static_cast<int>(diag_partials_needed(max_n));
    r |= alloc_into(c, c->diag_part, static_cast<size_t>(c->diag_part_cap));
    if (r != NBG_OK || hipHostMalloc(reinterpret_cast<void**>(&c->hsc), sizeof(DevScalars)) != hipSuccess) {
        nbg_destroy(c);
        return nullptr;
    }
    // the walk reads up to 8 particles past a leaf start; the pad stays zero (zero mass)
    if (hipMemset(c->spos, 0, (N + kSposPad) * sizeof(float4)) != hipSuccess ||
        hipMemset(c->rung, 0, N * sizeof(int32_t)) != hipSuccess ||
        hipMemset(c->dsc, 0, sizeof(DevScalars)) != hipSuccess) {
        nbg_destroy(c);
        return nullptr;
    }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            nbg_destroy(c);
            return nullptr;
        }
    return c;
}

void nbg_destroy(nbg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_dev(c->x); free_dev(c->y); free_dev(c->z); free_dev(c->m);
    free_dev(c->vx); free_dev(c->vy); free_dev(c->vz);
    free_dev(c->ax); free_dev(c->ay); free_dev(c->az);
    free_dev(c->backup); free_dev(c->pos_sys);
    free_dev(c->key_a); free_dev(c->key_b); free_dev(c->klo_a); free_dev(c->klo_b);
    free_dev(c->val_a); free_dev(c->val_b);
    free_dev(c->hist); free_dev(c->scan_part);
    free_dev(c->leaf_depth); free_dev(c->cpl_prev); free_dev(c->node_scan); free_dev(c->blk_depth); free_dev(c->leaf_list); free_dev(c->leaf_of);
    free_dev(c->spos);
    free_dev(c->stage4);
    free_dev(c->nd_start); free_dev(c->nd_count); free_dev(c->nd_depth); free_dev(c->nd_parent);
    free_dev(c->nd_oct); free_dev(c->nd_child); free_dev(c->nd_geom);
    free_dev(c->nd_rec); free_dev(c->nd_com64); free_dev(c->nd_rmax64);
    free_dev(c->depth_nodes); free_dev(c->dep_hist); free_dev(c->dep_off); free_dev(c->dep_part); free_dev(c->cnt_far); free_dev(c->cnt_near);
    free_dev(c->dsc); free_dev(c->diag_part);
    free_dev(c->aold); free_dev(c->rung); free_dev(c->tlist); free_dev(c->list_off); free_dev(c->red_part);
    free_dev(c->x64); free_dev(c->y64); free_dev(c->z64); free_dev(c->blk_bbox64); free_dev(c->srel);
    free_dev(c->nd_mom); free_dev(c->lacc); free_dev(c->lunit); free_dev(c->lpairs); free_dev(c->lpv); free_dev(c->lw32); free_dev(c->lw16);
    free_dev(c->adam_m); free_dev(c->adam_v); free_dev(c->pool_feat); free_dev(c->pool_tgt); free_dev(c->amag2);
    free_dev(c->grad_part); free_dev(c->loss_part); free_dev(c->loss_dev);
    free_dev(c->sys_off); free_dev(c->sys_root); free_dev(c->sys_geom); free_dev(c->blk_bbox);
    free_dev(c->blk_sys); free_dev(c->blk_start); free_dev(c->blk_count); free_dev(c->sys_blk0); free_dev(c->bin_start); free_dev(c->bin_count); free_dev(c->bin_cost);
    if (c->hsc) (void)hipHostFree(c->hsc);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* nbg_last_error(const nbg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int nbg_upload_batch(nbg_ctx* c, int32_t nsys, const int64_t* offsets, const float* x, const float* y, const float* z,
                     const float* m, const float* vx, const float* vy, const float* vz) {
    if (!c) return NBG_E_INVALID;
    if (nsys < 1 || nsys > kMaxSys || !offsets || offsets[0] != 0) {
        c->set_error("bad batch layout");
        return NBG_E_INVALID;
    }
    for (int s = 0; s < nsys; ++s)
        if (offsets[s + 1] < offsets[s]) {
            c->set_error("batch offsets must be non-decreasing");
            return NBG_E_INVALID;
        }
    const int64_t n = offsets[nsys];
    if (n < 0 || n > c->cap_n) {
        c->set_error("particle count exceeds the context capacity");
        return NBG_E_INVALID;
    }
    if (n > 0 && (!x || !y || !z || !m)) {
        c->set_error("positions and masses are required");
        return NBG_E_INVALID;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    drop_graph(c);
    c->f64_state = false;
    c->key_levels = kLevelsTiny;  // a new state tries the 13-level keys again
    c->n = static_cast<int>(n);
    c->nsys = nsys;
    c->h_sys_off.assign(offsets, offsets + nsys + 1);
    // sort blocks: slices of <= kSortTile keys inside one system
    std::vector<int32_t> bsys, bstart, bcnt, sb0(nsys + 1), off32(nsys + 1), psys(static_cast<size_t>(n));
    std::vector<int32_t> wstart, wcnt;
    for (int s = 0; s < nsys; ++s) {
        for (int64_t b = offsets[s]; b < offsets[s + 1]; b += 64) {
            wstart.push_back(static_cast<int32_t>(b));
            wcnt.push_back(static_cast<int32_t>(std::min<int64_t>(64, offsets[s + 1] - b)));
        }
        sb0[s] = static_cast<int32_t>(bsys.size());
        off32[s] = static_cast<int32_t>(offsets[s]);
        for (int64_t b = offsets[s]; b < offsets[s + 1]; b += kSortTile) {
            bsys.push_back(s);
            bstart.push_back(static_cast<int32_t>(b));
            bcnt.push_back(static_cast<int32_t>(std::min<int64_t>(kSortTile, offsets[s + 1] - b)));
        }
        for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i) psys[static_cast<size_t>(i)] = s;
    }
    sb0[nsys] = static_cast<int32_t>(bsys.size());
    off32[nsys] = static_cast<int32_t>(n);
    c->n_blocks = static_cast<int>(bsys.size());
    const size_t nb = std::max<size_t>(1, bsys.size());
    int r = NBG_OK;
    r |= alloc_into(c, c->blk_sys, nb);
    r |= alloc_into(c, c->blk_start, nb);
    r |= alloc_into(c, c->blk_count, nb);
    r |= alloc_into(c, c->sys_blk0, nsys + 1);
    r |= alloc_into(c, c->sys_off, nsys + 1);
    r |= alloc_into(c, c->sys_root, nsys);
    r |= alloc_into(c, c->sys_geom, 4 * static_cast<size_t>(nsys));
    r |= alloc_into(c, c->blk_bbox, 6 * nb);
    r |= alloc_into(c, c->hist, static_cast<size_t>(kRadix) * nb);
    c->n_bins = static_cast<int>(wstart.size());
    r |= alloc_into(c, c->bin_start, std::max<size_t>(1, wstart.size()));
    r |= alloc_into(c, c->bin_count, std::max<size_t>(1, wcnt.size()));
    r |= alloc_into(c, c->bin_cost, std::max<size_t>(1, wcnt.size()));
    // (also k_leaf_depth's per-block sums, one per 256 positions)
    const size_t parts = std::max<size_t>(std::max<size_t>(grid_for(n + 1, 2048), grid_for(n, 256)),
                                          grid_for(static_cast<int64_t>(kRadix) * nb, 2048)) + 8;
    r |= alloc_into(c, c->scan_part, parts);
    if (r != NBG_OK) return r;
    if (!bsys.empty()) {
        NBG_HIP_CHECK(c, hipMemcpy(c->blk_sys, bsys.data(), bsys.size() * 4, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->blk_start, bstart.data(), bstart.size() * 4, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->blk_count, bcnt.data(), bcnt.size() * 4, hipMemcpyHostToDevice));
    }
    NBG_HIP_CHECK(c, hipMemcpy(c->sys_blk0, sb0.data(), sb0.size() * 4, hipMemcpyHostToDevice));
    if (!wstart.empty()) {
        NBG_HIP_CHECK(c, hipMemcpy(c->bin_start, wstart.data(), wstart.size() * 4, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->bin_count, wcnt.data(), wcnt.size() * 4, hipMemcpyHostToDevice));
    }
    NBG_HIP_CHECK(c, hipMemcpy(c->sys_off, off32.data(), off32.size() * 4, hipMemcpyHostToDevice));
    // no cost history for the new layout
    NBG_HIP_CHECK(c, hipMemset(c->bin_cost, 0, std::max<size_t>(1, wcnt.size()) * 4));
    NBG_HIP_CHECK(c, hipMemset(&c->dsc->pp_cost_sum, 0, 2 * sizeof(uint64_t)));
    if (n > 0) {
        NBG_HIP_CHECK(c, hipMemcpy(c->pos_sys, psys.data(), psys.size() * 4, hipMemcpyHostToDevice));
        const size_t bytes = static_cast<size_t>(n) * sizeof(float);
        NBG_HIP_CHECK(c, hipMemcpy(c->x, x, bytes, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->y, y, bytes, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->z, z, bytes, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->m, m, bytes, hipMemcpyHostToDevice));
        float* vdst[3] = {c->vx, c->vy, c->vz};
        const float* vsrc[3] = {vx, vy, vz};
        for (int k = 0; k < 3; ++k) {
            if (vsrc[k])
                NBG_HIP_CHECK(c, hipMemcpy(vdst[k], vsrc[k], bytes, hipMemcpyHostToDevice));
            else
                NBG_HIP_CHECK(c, hipMemset(vdst[k], 0, bytes));
        }
        NBG_HIP_CHECK(c, hipMemset(c->ax, 0, bytes));
        NBG_HIP_CHECK(c, hipMemset(c->ay, 0, bytes));
        NBG_HIP_CHECK(c, hipMemset(c->az, 0, bytes));
    }
    if (n > 0) NBG_HIP_CHECK(c, hipMemset(c->rung, 0, static_cast<size_t>(n) * sizeof(int32_t)));
    NBG_HIP_CHECK(c, hipDeviceSynchronize());
    c->have_state = true;
    c->have_tree = false;
    c->have_rungs = false;
    return NBG_OK;
}

int nbg_upload(nbg_ctx* c, int64_t n, const float* x, const float* y, const float* z, const float* m,
               const float* vx, const float* vy, const float* vz) {
    const int64_t off[2] = {0, n};
    return nbg_upload_batch(c, 1, off, x, y, z, m, vx, vy, vz);
}

int nbg_upload_f64(nbg_ctx* c, int64_t n, const double* x, const double* y, const double* z, const float* m,
                   const float* vx, const float* vy, const float* vz) {
    if (!c) return NBG_E_INVALID;
    if (n < 0 || n > c->cap_n) {
        c->set_error("particle count exceeds the context capacity");
        return NBG_E_INVALID;
    }
    if (n > 0 && (!x || !y || !z || !m)) {
        c->set_error("positions and masses are required");
        return NBG_E_INVALID;
    }
    // the FP32 roundings serve every kernel that does not take the FP64 positions
    std::vector<float> fx(static_cast<size_t>(n)), fy(static_cast<size_t>(n)), fz(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        fx[static_cast<size_t>(i)] = static_cast<float>(x[i]);
        fy[static_cast<size_t>(i)] = static_cast<float>(y[i]);
        fz[static_cast<size_t>(i)] = static_cast<float>(z[i]);
    }
    if (int r = nbg_upload(c, n, fx.data(), fy.data(), fz.data(), m, vx, vy, vz)) return r;
    const size_t cap = static_cast<size_t>(c->cap_n);
    int r = NBG_OK;
    if (!c->x64) r |= alloc_into(c, c->x64, cap);
    if (!c->y64) r |= alloc_into(c, c->y64, cap);
    if (!c->z64) r |= alloc_into(c, c->z64, cap);
    r |= alloc_into(c, c->blk_bbox64, 6 * static_cast<size_t>(std::max(1, c->n_blocks)));
    if (!c->srel) {
        r |= alloc_into(c, c->srel, cap + kSposPad);
        // the zero pad after the records (the near field loads whole groups); entries past
        // n stay finite as n changes
        if (c->srel) NBG_HIP_CHECK(c, hipMemset(c->srel, 0, (cap + kSposPad) * sizeof(float4)));
    }
    if (r != NBG_OK) return r;
    if (n > 0) {
        const size_t bytes = static_cast<size_t>(n) * sizeof(double);
        NBG_HIP_CHECK(c, hipMemcpy(c->x64, x, bytes, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->y64, y, bytes, hipMemcpyHostToDevice));
        NBG_HIP_CHECK(c, hipMemcpy(c->z64, z, bytes, hipMemcpyHostToDevice));
    }
    c->f64_state = true;
    return NBG_OK;
}

int nbg_upload_accel(nbg_ctx* c, const float* ax, const float* ay, const float* az) {
    if (!c) return NBG_E_INVALID;
    if (!c->have_state) {
        c->set_error("no particles uploaded");
        return NBG_E_STATE;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    const size_t bytes = static_cast<size_t>(c->n) * sizeof(float);
    if (bytes == 0) return NBG_OK;
    float* dst[3] = {c->ax, c->ay, c->az};
    const float* src[3] = {ax, ay, az};
    for (int k = 0; k < 3; ++k) {
        if (src[k])
            NBG_HIP_CHECK(c, hipMemcpy(dst[k], src[k], bytes, hipMemcpyHostToDevice));
        else
            NBG_HIP_CHECK(c, hipMemset(dst[k], 0, bytes));
    }
    return NBG_OK;
}

int nbg_accel(nbg_ctx* c, const nbg_params* p, nbg_stats* stats) {
    if (!c) return NBG_E_INVALID;
    if (int r = validate(c, p)) return r;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    if (c->n == 0) {
        if (stats) std::memset(stats, 0, sizeof(*stats));
        return NBG_OK;
    }
    return accel_eager(c, *p, stats);
}

int nbg_step(nbg_ctx* c, const nbg_params* p, int32_t nsteps, nbg_stats* stats) {
    if (!c) return NBG_E_INVALID;
    if (int r = validate(c, p)) return r;
    if (nsteps < 0) {
        c->set_error("nsteps must be >= 0");
        return NBG_E_INVALID;
    }
    if (c->f64_state) {
        c->set_error("an FP64 state (nbg_upload_f64) takes force evaluations only");
        return NBG_E_STATE;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    if (c->n == 0 || nsteps == 0) {
        if (stats) std::memset(stats, 0, sizeof(*stats));
        return NBG_OK;
    }
    if (stats) std::memset(stats, 0, sizeof(*stats));
    c->last_G = p->G;
    NBG_HIP_CHECK(c, hipEventRecord(c->ev[6], c->stream));
    if (p->block) {
        if (int r = block_steps(c, *p, nsteps, stats)) return r;
        NBG_HIP_CHECK(c, hipEventRecord(c->ev[7], c->stream));
        NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        if (stats) {
            if (p->solver == NBG_SOLVER_BARNES_HUT) {
                if (int r = read_scalars(c)) return r;
                fill_tree_stats(c, stats);
            }
            stats->ms_total = elapsed(c->ev[6], c->ev[7]);
        }
        return NBG_OK;
    }
    if (p->adaptive) NBG_HIP_CHECK(c, hipMemsetAsync(&c->dsc->t_accum, 0, sizeof(double), c->stream));
    const bool use_graph = !(p->flags & NBG_FLAG_NO_GRAPH) && c->last_levels <= kLevelsShort;
    if (!use_graph) {
        if (int r = steps_eager(c, *p, nsteps, stats)) return r;
    } else {
        if (!c->graph_valid || !same_params(c->graph_params, *p) || c->graph_levels != c->last_levels) {
            drop_graph(c);
            (void)hipGetLastError();
            NBG_HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            c->capturing = true;
            enqueue_step(c, *p, c->last_levels, (p->flags & NBG_FLAG_STAGE_TIMES) != 0);
            c->capturing = false;
            const hipError_t le = hipGetLastError();
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(c->stream, &g);
            if (le != hipSuccess || ee != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                c->set_error(std::string("graph capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ee));
                return NBG_E_HIP;
            }
            c->graph = g;
            NBG_HIP_CHECK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
            c->graph_params = *p;
            c->graph_levels = c->last_levels;
            c->graph_valid = true;
        }
        if (int r = save_backup(c)) return r;
        NBG_HIP_CHECK(c, hipMemsetAsync(&c->dsc->sticky_overflow, 0, 2 * sizeof(int32_t), c->stream));
        NBG_HIP_CHECK(c, hipMemsetAsync(&c->dsc->sticky_learn, 0, sizeof(int32_t), c->stream));
        const bool per_step = (p->flags & NBG_FLAG_STAGE_TIMES) && stats;
        for (int s = 0; s < nsteps; ++s) {
            NBG_HIP_CHECK(c, hipGraphLaunch(c->graph_exec, c->stream));
            if (per_step) {
                NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
                add_stage_times(c, *p, stats);
            }
        }
        NBG_HIP_CHECK(c, hipMemcpyAsync(&c->hsc->sticky_overflow, &c->dsc->sticky_overflow, 2 * sizeof(int32_t),
                                        hipMemcpyDeviceToHost, c->stream));
        NBG_HIP_CHECK(c, hipMemcpyAsync(&c->hsc->sticky_learn, &c->dsc->sticky_learn, sizeof(int32_t),
                                        hipMemcpyDeviceToHost, c->stream));
        NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        const bool overflow = c->hsc->sticky_overflow != 0 || c->hsc->sticky_learn != 0;
        const bool cap = c->hsc->sticky_cap != 0;
        if (cap) {
            // steps past the truncated tree are discarded: the caller gets the state it passed in
            if (int r = restore_backup(c)) return r;
            c->have_tree = false;
            c->set_error("node pool capacity exceeded");
            return NBG_E_CAPACITY;
        }
        if (overflow) {
            if (int r = restore_backup(c)) return r;
            if (stats) std::memset(stats, 0, sizeof(*stats));
            if (p->adaptive) NBG_HIP_CHECK(c, hipMemsetAsync(&c->dsc->t_accum, 0, sizeof(double), c->stream));
            if (int r = steps_eager(c, *p, nsteps, stats)) return r;
        }
    }
    NBG_HIP_CHECK(c, hipEventRecord(c->ev[7], c->stream));
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->have_tree = p->solver == NBG_SOLVER_BARNES_HUT;
    if (stats) {
        if (int r = read_scalars(c)) return r;
        fill_tree_stats(c, stats);
        stats->ms_total = elapsed(c->ev[6], c->ev[7]);
        stats->particle_evals = static_cast<int64_t>(nsteps) * c->n;
        if (p->adaptive) {
            stats->t_advanced = c->hsc->t_accum;
            stats->dt_last = c->hsc->dt_step;
        } else {
            stats->t_advanced = p->dt * nsteps;
            stats->dt_last = p->dt;
        }
    }
    return NBG_OK;
}

int nbg_download(nbg_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* ax, float* ay,
                 float* az) {
    if (!c) return NBG_E_INVALID;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const size_t bytes = static_cast<size_t>(c->n) * sizeof(float);
    if (bytes == 0) return NBG_OK;
    float* dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    const float* src[9] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az};
    for (int k = 0; k < 9; ++k)
        if (dst[k]) NBG_HIP_CHECK(c, hipMemcpy(dst[k], src[k], bytes, hipMemcpyDeviceToHost));
    return NBG_OK;
}

int nbg_tree_export(nbg_ctx* c, int64_t* n_nodes, uint64_t* keys, int32_t* order, double* center, int32_t* topo,
                    int32_t* children, float* moments) {
    if (!c || !n_nodes) return NBG_E_INVALID;
    if (!c->have_tree) {
        c->set_error("no Barnes-Hut tree: run nbg_accel with the Barnes-Hut solver first");
        return NBG_E_STATE;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    if (int r = read_scalars(c)) return r;
    const int64_t nn = c->hsc->n_nodes;
    const int64_t cap = *n_nodes;
    *n_nodes = nn;
    const size_t n = static_cast<size_t>(c->n);
    if (keys) NBG_HIP_CHECK(c, hipMemcpy(keys, c->key_sorted, n * 8, hipMemcpyDeviceToHost));
    if (order) NBG_HIP_CHECK(c, hipMemcpy(order, c->perm, n * 4, hipMemcpyDeviceToHost));
    if (!center && !topo && !children && !moments) return NBG_OK;
    if (cap < nn) {
        c->set_error("node arrays too small");
        return NBG_E_INVALID;
    }
    const size_t un = static_cast<size_t>(nn);
    if (center) NBG_HIP_CHECK(c, hipMemcpy(center, c->nd_geom, un * 4 * sizeof(double), hipMemcpyDeviceToHost));
    if (children) NBG_HIP_CHECK(c, hipMemcpy(children, c->nd_child, un * 8 * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<float4> aux(un);
    NBG_HIP_CHECK(c, copy_rec_slot(c, aux.data(), 3, nn));
    if (topo) {
        std::vector<int32_t> st(un), cnt(un), dep(un), par(un);
        NBG_HIP_CHECK(c, hipMemcpy(st.data(), c->nd_start, un * 4, hipMemcpyDeviceToHost));
        NBG_HIP_CHECK(c, hipMemcpy(cnt.data(), c->nd_count, un * 4, hipMemcpyDeviceToHost));
        NBG_HIP_CHECK(c, hipMemcpy(dep.data(), c->nd_depth, un * 4, hipMemcpyDeviceToHost));
        NBG_HIP_CHECK(c, hipMemcpy(par.data(), c->nd_parent, un * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < un; ++i) {
            uint32_t info;
            std::memcpy(&info, &aux[i].w, 4);
            const bool leaf = (info & kLeafBit) != 0;
            topo[5 * i + 0] = par[i];
            topo[5 * i + 1] = leaf ? st[i] : -1;
            topo[5 * i + 2] = leaf ? cnt[i] : 0;
            topo[5 * i + 3] = dep[i];
            topo[5 * i + 4] = leaf ? 0 : static_cast<int32_t>(info);
        }
    }
    if (moments) {
        std::vector<float4> com(un), quad(un);
        std::vector<double4> com64(un);
        NBG_HIP_CHECK(c, hipMemcpy(com64.data(), c->nd_com64, un * sizeof(double4), hipMemcpyDeviceToHost));
        NBG_HIP_CHECK(c, copy_rec_slot(c, com.data(), 0, nn));
        NBG_HIP_CHECK(c, copy_rec_slot(c, quad.data(), 1, nn));
        // device masses and quadrupoles carry G; report them without it
        const float ig = (c->last_G != 0.0) ? static_cast<float>(1.0 / c->last_G) : 1.0f;
        for (size_t i = 0; i < un; ++i) {
            float* o = moments + 10 * i;
            o[0] = com[i].x;
            o[1] = com[i].y;
            o[2] = com[i].z;
            o[3] = static_cast<float>(com64[i].w);
            o[4] = aux[i].z;
            o[5] = quad[i].x * ig;
            o[6] = quad[i].y * ig;
            o[7] = quad[i].z * ig;
            o[8] = quad[i].w * ig;
            o[9] = aux[i].x * ig;
        }
    }
    return NBG_OK;
}

int nbg_interaction_counts(nbg_ctx* c, int64_t n_nodes, int64_t* n_far, int64_t* n_near) {
    if (!c || !n_far || !n_near) return NBG_E_INVALID;
    if (!c->have_tree) {
        c->set_error("no Barnes-Hut tree");
        return NBG_E_STATE;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    if (int r = read_scalars(c)) return r;
    const int64_t nn = c->hsc->n_nodes;
    if (n_nodes < nn) {
        c->set_error("count arrays too small");
        return NBG_E_INVALID;
    }
    std::vector<int32_t> f(static_cast<size_t>(nn)), q(static_cast<size_t>(nn));
    NBG_HIP_CHECK(c, hipMemcpy(f.data(), c->cnt_far, f.size() * 4, hipMemcpyDeviceToHost));
    NBG_HIP_CHECK(c, hipMemcpy(q.data(), c->cnt_near, q.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nn; ++i) {
        n_far[i] = f[static_cast<size_t>(i)];
        n_near[i] = q[static_cast<size_t>(i)];
    }
    return NBG_OK;
}

int nbg_diagnostics(nbg_ctx* c, const nbg_params* p, int32_t with_potential, double* out) {
    if (!c || !out) return NBG_E_INVALID;
    if (int r = validate(c, p)) return r;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    int r = diagnostics(c, *p, with_potential, out, 0);
    if (r == NBG_OK) r = check_launch(c);
    return r;
}

int nbg_accel_direct_targets(nbg_ctx* c, const nbg_params* p, int64_t nt, const int64_t* targets, float* ax,
                             float* ay, float* az) {
    if (!c) return NBG_E_INVALID;
    if (int r = validate(c, p)) return r;
    if (nt < 0 || (nt > 0 && (!targets || !ax || !ay || !az))) {
        c->set_error("bad target list");
        return NBG_E_INVALID;
    }
    if (nt == 0) return NBG_OK;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    std::vector<int32_t> t32(static_cast<size_t>(nt));
    for (int64_t k = 0; k < nt; ++k) {
        if (targets[k] < 0 || targets[k] >= c->n) {
            c->set_error("target index out of range");
            return NBG_E_INVALID;
        }
        t32[static_cast<size_t>(k)] = static_cast<int32_t>(targets[k]);
    }
    int32_t* dt = nullptr;
    float* dout = nullptr;
    NBG_HIP_CHECK(c, hipMalloc(&dt, t32.size() * 4));
    if (hipMalloc(&dout, t32.size() * 12) != hipSuccess) {
        (void)hipFree(dt);
        c->set_error("hipMalloc failed");
        return NBG_E_OOM;
    }
    int r = NBG_OK;
    if (hipMemcpy(dt, t32.data(), t32.size() * 4, hipMemcpyHostToDevice) != hipSuccess) r = NBG_E_HIP;
    if (r == NBG_OK) {
        launch_direct_targets(c, *p, nt, dt, nullptr, dout);
        r = check_launch(c);
    }
    std::vector<float> h(static_cast<size_t>(nt) * 3);
    if (r == NBG_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = NBG_E_HIP;
    if (r == NBG_OK && hipMemcpy(h.data(), dout, h.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) r = NBG_E_HIP;
    (void)hipFree(dt);
    (void)hipFree(dout);
    if (r != NBG_OK) {
        if (c->err.empty()) c->set_error("direct-targets evaluation failed");
        return r;
    }
    for (int64_t k = 0; k < nt; ++k) {
        ax[k] = h[3 * static_cast<size_t>(k) + 0];
        ay[k] = h[3 * static_cast<size_t>(k) + 1];
        az[k] = h[3 * static_cast<size_t>(k) + 2];
    }
    c->have_tree = false;
    return NBG_OK;
}

int nbg_accel_bh_targets(nbg_ctx* c, const nbg_params* p, int64_t nt, const int64_t* targets, float* ax, float* ay,
                         float* az) {
    if (!c) return NBG_E_INVALID;
    if (int r = validate(c, p)) return r;
    if (p->solver != NBG_SOLVER_BARNES_HUT || (p->flags & NBG_FLAG_LEARNED)) {
        c->set_error("nbg_accel_bh_targets needs the Barnes-Hut solver without the learned correction");
        return NBG_E_INVALID;
    }
    if (nt < 0 || nt > c->cap_n || (nt > 0 && (!targets || !ax || !ay || !az))) {
        c->set_error("bad target list");
        return NBG_E_INVALID;
    }
    for (int64_t k = 0; k < nt; ++k)
        if (targets[k] < 0 || targets[k] >= c->n) {
            c->set_error("target index out of range");
            return NBG_E_INVALID;
        }
    if (nt == 0) return NBG_OK;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    nbg_params q = *p;
    q.mac = NBG_MAC_GEOMETRIC;
    q.flags &= ~NBG_FLAG_COUNT;
    if (int r = build_tree_eager(c, q)) return r;
    // original index -> sorted position
    const size_t n = static_cast<size_t>(c->n);
    std::vector<uint32_t> perm(n);
    std::vector<int32_t> inv(n), list(static_cast<size_t>(nt));
    NBG_HIP_CHECK(c, hipMemcpy(perm.data(), c->perm, n * 4, hipMemcpyDeviceToHost));
    for (size_t pos = 0; pos < n; ++pos) inv[perm[pos]] = static_cast<int32_t>(pos);
    for (int64_t k = 0; k < nt; ++k) list[static_cast<size_t>(k)] = inv[static_cast<size_t>(targets[k])];
    const uint32_t cnt = static_cast<uint32_t>(nt);
    NBG_HIP_CHECK(c, hipMemcpy(c->tlist, list.data(), list.size() * 4, hipMemcpyHostToDevice));
    NBG_HIP_CHECK(c, hipMemcpy(&c->dsc->tcount, &cnt, sizeof(cnt), hipMemcpyHostToDevice));
    float* dout = nullptr;
    NBG_HIP_CHECK(c, hipMalloc(&dout, list.size() * 12));
    launch_walk_pp(c, q, c->tlist, &c->dsc->tcount, dout);
    int r = check_launch(c);
    std::vector<float> h(list.size() * 3);
    if (r == NBG_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = NBG_E_HIP;
    if (r == NBG_OK && hipMemcpy(h.data(), dout, h.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) r = NBG_E_HIP;
    (void)hipFree(dout);
    if (r != NBG_OK) {
        if (c->err.empty()) c->set_error("Barnes-Hut targets evaluation failed");
        return r;
    }
    for (int64_t k = 0; k < nt; ++k) {
        ax[k] = h[3 * static_cast<size_t>(k) + 0];
        ay[k] = h[3 * static_cast<size_t>(k) + 1];
        az[k] = h[3 * static_cast<size_t>(k) + 2];
    }
    return NBG_OK;
}

int nbg_rungs(nbg_ctx* c, int32_t* rungs) {
    if (!c || !rungs) return NBG_E_INVALID;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (c->n > 0)
        NBG_HIP_CHECK(c, hipMemcpy(rungs, c->rung, static_cast<size_t>(c->n) * 4, hipMemcpyDeviceToHost));
    return NBG_OK;
}

// ---------------------------------------------------------------------------- learned correction

extern "C++" {
namespace {
uint16_t host_bf16(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    b += 0x7fffu + ((b >> 16) & 1u);
    return static_cast<uint16_t>(b >> 16);
}

template <typename T>
int grow_copy(nbg_ctx* c, T*& p, size_t old_count, size_t new_count) {
    T* q = nullptr;
    hipError_t e;
    q = dev_alloc<T>(new_count, e);
    if (!q) {
        c->set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return NBG_E_OOM;
    }
    if (p && old_count) NBG_HIP_CHECK(c, hipMemcpy(q, p, old_count * sizeof(T), hipMemcpyDeviceToDevice));
    if (p) (void)hipFree(p);
    p = q;
    return NBG_OK;
}
}  // namespace
}  // extern "C++"

int nbg_learn_set_weights(nbg_ctx* c, const float* blob, int64_t n) {
    if (!c) return NBG_E_INVALID;
    if (!blob || n != NBG_LEARN_NPARAM) {
        c->set_error("model blob must hold NBG_LEARN_NPARAM floats");
        return NBG_E_INVALID;
    }
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(blob[i])) {
            c->set_error("model blob holds a non-finite value");
            return NBG_E_INVALID;
        }
    const float act = blob[NBG_LEARN_ACT_SLOT];
    if (act != static_cast<float>(NBG_LEARN_ACT_SILU) && act != static_cast<float>(NBG_LEARN_ACT_RELU)) {
        c->set_error("model blob: unknown activation code in blob[NBG_LEARN_ACT_SLOT]");
        return NBG_E_INVALID;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    int r = NBG_OK;
    if (!c->lw32) {
        r |= alloc_into(c, c->lw32, NBG_LEARN_NPARAM);
        r |= alloc_into(c, c->lw16, NBG_LEARN_NTRAIN + 8);
        r |= alloc_into(c, c->adam_m, NBG_LEARN_NTRAIN);
        r |= alloc_into(c, c->adam_v, NBG_LEARN_NTRAIN);
        if (r != NBG_OK) return r;
    }
    std::vector<uint16_t> h16(NBG_LEARN_NTRAIN + 8, 0);
    for (int i = 0; i < NBG_LEARN_NTRAIN; ++i) h16[static_cast<size_t>(i)] = host_bf16(blob[32 + i]);
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    NBG_HIP_CHECK(c, hipMemcpy(c->lw32, blob, NBG_LEARN_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    NBG_HIP_CHECK(c, hipMemcpy(c->lw16, h16.data(), h16.size() * 2, hipMemcpyHostToDevice));
    NBG_HIP_CHECK(c, hipMemset(c->adam_m, 0, NBG_LEARN_NTRAIN * sizeof(float)));
    NBG_HIP_CHECK(c, hipMemset(c->adam_v, 0, NBG_LEARN_NTRAIN * sizeof(float)));
    c->adam_t = 0;
    c->learn_act = static_cast<int>(act);
    c->have_weights = true;
    return NBG_OK;
}

int nbg_learn_get_weights(nbg_ctx* c, float* blob, int64_t n) {
    if (!c) return NBG_E_INVALID;
    if (!blob || n < NBG_LEARN_NPARAM) {
        c->set_error("blob must hold NBG_LEARN_NPARAM floats");
        return NBG_E_INVALID;
    }
    if (!c->have_weights) {
        c->set_error("no learned model installed");
        return NBG_E_STATE;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    NBG_HIP_CHECK(c, hipMemcpy(blob, c->lw32, NBG_LEARN_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return NBG_OK;
}

int nbg_learn_clear(nbg_ctx* c) {
    if (!c) return NBG_E_INVALID;
    c->pool_n = 0;
    return NBG_OK;
}

int nbg_learn_collect(nbg_ctx* c, const nbg_params* p, int64_t max_new, int64_t* n_pool) {
    if (!c) return NBG_E_INVALID;
    if (!p) {
        c->set_error("params is NULL");
        return NBG_E_INVALID;
    }
    nbg_params q = *p;
    q.solver = NBG_SOLVER_BARNES_HUT;
    q.multipole = NBG_QUADRUPOLE;
    q.flags = NBG_FLAG_KAHAN;
    if (int r = validate(c, &q)) return r;
    if (!(q.learn_band >= 0.0)) {
        c->set_error("learn_band must be >= 0");
        return NBG_E_INVALID;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    if (int r = learn_alloc(c)) return r;
    if (c->n == 0) {
        if (n_pool) *n_pool = c->pool_n;
        return NBG_OK;
    }
    // the direct sum and the listing walk below overwrite the context's accelerations:
    // they are restored at the end, so a(t) of the caller's last evaluation survives
    if (int r = save_backup(c)) return r;
    // direct-sum accelerations of the snapshot: the loss weights use |a|
    nbg_params d = q;
    d.solver = NBG_SOLVER_DIRECT;
    c->last_G = q.G;
    launch_direct(c, d);
    launch_learn_amag2(c);
    if (int r = check_launch(c)) return r;
    // the walk with the pair list (no correction applied)
    q.flags |= NBG_FLAG_LEARNED;
    if (int r = accel_eager(c, q, nullptr, true)) return r;
    const int64_t npairs = c->hsc->learn_pairs;
    const int64_t nnew = (max_new > 0) ? std::min(max_new, npairs) : npairs;
    if (c->pool_n + nnew > c->pool_cap) {
        const int64_t cap = std::max<int64_t>(c->pool_n + nnew, c->pool_cap + c->pool_cap / 2);
        if (int r = grow_copy(c, c->pool_feat, static_cast<size_t>(16 * c->pool_n), static_cast<size_t>(16 * cap)))
            return r;
        if (int r = grow_copy(c, c->pool_tgt, static_cast<size_t>(4 * c->pool_n), static_cast<size_t>(4 * cap)))
            return r;
        c->pool_cap = cap;
    }
    if (nnew > 0) launch_learn_samples(c, q, npairs, nnew, c->pool_n);
    if (int r = restore_backup(c)) return r;
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->pool_n += nnew;
    c->have_tree = true;
    if (n_pool) *n_pool = c->pool_n;
    return NBG_OK;
}

int nbg_learn_pool(nbg_ctx* c, int64_t offset, int64_t count, float* feats, float* targets) {
    if (!c) return NBG_E_INVALID;
    if (offset < 0 || count < 0 || offset + count > c->pool_n) {
        c->set_error("pool range out of bounds");
        return NBG_E_INVALID;
    }
    if (count == 0) return NBG_OK;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (feats)
        NBG_HIP_CHECK(c, hipMemcpy(feats, c->pool_feat + 16 * offset, static_cast<size_t>(count) * 64,
                                   hipMemcpyDeviceToHost));
    if (targets)
        NBG_HIP_CHECK(c, hipMemcpy(targets, c->pool_tgt + 4 * offset, static_cast<size_t>(count) * 16,
                                   hipMemcpyDeviceToHost));
    return NBG_OK;
}

int nbg_learn_standardize(nbg_ctx* c, int64_t count) {
    if (!c) return NBG_E_INVALID;
    if (!c->have_weights) {
        c->set_error("no learned model installed");
        return NBG_E_STATE;
    }
    const int64_t n = (count <= 0 || count > c->pool_n) ? c->pool_n : count;
    if (n < 2) {
        c->set_error("the pool holds fewer than two samples");
        return NBG_E_STATE;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    double* dsum = nullptr;
    NBG_HIP_CHECK(c, hipMalloc(&dsum, 33 * sizeof(double)));
    int r = NBG_OK;
    if (hipMemsetAsync(dsum, 0, 33 * sizeof(double), c->stream) != hipSuccess) r = NBG_E_HIP;
    double h[33] = {};
    if (r == NBG_OK) {
        launch_learn_stats(c, n, dsum);
        r = check_launch(c);
    }
    if (r == NBG_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = NBG_E_HIP;
    if (r == NBG_OK && hipMemcpy(h, dsum, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) r = NBG_E_HIP;
    (void)hipFree(dsum);
    if (r != NBG_OK) {
        c->set_error("pool statistics failed");
        return r;
    }
    float norm[32];
    for (int k = 0; k < 16; ++k) {
        const double mu = h[k] / static_cast<double>(n);
        const double var = std::max(0.0, h[16 + k] / static_cast<double>(n) - mu * mu);
        const bool used = k < NBG_LEARN_FEATURES && var > 1e-20;
        norm[k] = used ? static_cast<float>(mu) : 0.f;
        norm[16 + k] = used ? static_cast<float>(1.0 / std::sqrt(var)) : 1.f;
    }
    norm[NBG_LEARN_ACT_SLOT] = static_cast<float>(c->learn_act);  // the model's activation code stays
    const double wmean = h[32] / static_cast<double>(n);
    c->pool_inv_wmean = (wmean > 0.0) ? static_cast<float>(1.0 / wmean) : 1.f;
    NBG_HIP_CHECK(c, hipMemcpy(c->lw32, norm, sizeof(norm), hipMemcpyHostToDevice));
    return NBG_OK;
}

extern "C++" {
namespace {
int train_buffers(nbg_ctx* c, int nblocks, int nloss) {
    if (nblocks > c->grad_blocks_cap) {
        if (int r = alloc_into(c, c->grad_part, static_cast<size_t>(nblocks) * NBG_LEARN_NTRAIN)) return r;
        if (int r = alloc_into(c, c->loss_part, static_cast<size_t>(nblocks))) return r;
        c->grad_blocks_cap = nblocks;
    }
    if (nloss > c->loss_cap) {
        if (int r = alloc_into(c, c->loss_dev, static_cast<size_t>(nloss))) return r;
        c->loss_cap = nloss;
    }
    return NBG_OK;
}
}  // namespace
}  // extern "C++"

int nbg_learn_train(nbg_ctx* c, int32_t steps, int32_t batch, float lr, uint64_t seed, float* loss) {
    if (!c) return NBG_E_INVALID;
    if (!c->have_weights) {
        c->set_error("no learned model installed");
        return NBG_E_STATE;
    }
    if (steps < 0 || batch < 64 || batch % 64 != 0 || batch > (1 << 22) || !(lr > 0.f)) {
        c->set_error("steps >= 0, batch a multiple of 64 (<= 4M) and lr > 0 are required");
        return NBG_E_INVALID;
    }
    if (c->pool_n < 1) {
        c->set_error("the sample pool is empty (nbg_learn_collect)");
        return NBG_E_STATE;
    }
    if (steps == 0) return NBG_OK;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    const int nblocks = batch / 64;
    if (int r = train_buffers(c, nblocks, steps)) return r;
    for (int s = 0; s < steps; ++s) {
        launch_learn_fwdbwd(c, c->pool_feat, c->pool_tgt, c->pool_n, nblocks, seed, static_cast<int>(c->adam_t), 1,
                            c->pool_inv_wmean, 1.f / static_cast<float>(batch), nullptr);
        ++c->adam_t;
        launch_learn_adam(c, nblocks, lr, c->adam_t, c->loss_dev + s);
        if (int r = check_launch(c)) return r;
    }
    NBG_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (loss) {
        NBG_HIP_CHECK(c, hipMemcpy(loss, c->loss_dev, static_cast<size_t>(steps) * 4, hipMemcpyDeviceToHost));
        for (int s = 0; s < steps; ++s) loss[s] /= static_cast<float>(batch);
    }
    return NBG_OK;
}

int nbg_learn_grad(nbg_ctx* c, int64_t n, const float* feats, const float* targets, float* out, float* loss,
                   float* grad) {
    if (!c) return NBG_E_INVALID;
    if (!c->have_weights) {
        c->set_error("no learned model installed");
        return NBG_E_STATE;
    }
    if (n < 1 || n > (1 << 22) || !feats || !targets) {
        c->set_error("1 <= n <= 4M samples with features and targets are required");
        return NBG_E_INVALID;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    const int nblocks = static_cast<int>((n + 63) / 64);
    if (int r = train_buffers(c, nblocks, 1)) return r;
    float *df = nullptr, *dt = nullptr, *dout = nullptr, *dgrad = nullptr;
    int r = NBG_OK;
    if (hipMalloc(&df, static_cast<size_t>(n) * 64) != hipSuccess) r = NBG_E_OOM;
    if (r == NBG_OK && hipMalloc(&dt, static_cast<size_t>(n) * 16) != hipSuccess) r = NBG_E_OOM;
    if (r == NBG_OK && hipMalloc(&dout, static_cast<size_t>(n) * 12) != hipSuccess) r = NBG_E_OOM;
    if (r == NBG_OK && hipMalloc(&dgrad, static_cast<size_t>(NBG_LEARN_NTRAIN + 1) * 4) != hipSuccess) r = NBG_E_OOM;
    if (r == NBG_OK && hipMemcpy(df, feats, static_cast<size_t>(n) * 64, hipMemcpyHostToDevice) != hipSuccess)
        r = NBG_E_HIP;
    if (r == NBG_OK && hipMemcpy(dt, targets, static_cast<size_t>(n) * 16, hipMemcpyHostToDevice) != hipSuccess)
        r = NBG_E_HIP;
    if (r == NBG_OK) {
        launch_learn_fwdbwd(c, df, dt, n, nblocks, 0, 0, 0, 1.f, 1.f / static_cast<float>(n), dout);
        launch_learn_reduce(c, nblocks, dgrad, dgrad + NBG_LEARN_NTRAIN);
        r = check_launch(c);
    }
    if (r == NBG_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = NBG_E_HIP;
    float lsum = 0.f;
    if (r == NBG_OK && out && hipMemcpy(out, dout, static_cast<size_t>(n) * 12, hipMemcpyDeviceToHost) != hipSuccess)
        r = NBG_E_HIP;
    if (r == NBG_OK && grad &&
        hipMemcpy(grad, dgrad, NBG_LEARN_NTRAIN * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        r = NBG_E_HIP;
    if (r == NBG_OK && hipMemcpy(&lsum, dgrad + NBG_LEARN_NTRAIN, 4, hipMemcpyDeviceToHost) != hipSuccess)
        r = NBG_E_HIP;
    (void)hipFree(df);
    (void)hipFree(dt);
    (void)hipFree(dout);
    (void)hipFree(dgrad);
    if (r != NBG_OK) {
        if (c->err.empty()) c->set_error("learned forward/backward failed");
        return r;
    }
    if (loss) *loss = lsum / static_cast<float>(n);
    return NBG_OK;
}

// ---------------------------------------------------------------------------- tuning

namespace {
int check_tuning(nbg_ctx* c, const nbg_tuning& t) {
    const bool ok = (t.walk_kernel == 0 || t.walk_kernel == 1) &&
                    (t.near_tile == 4 || t.near_tile == 8 || t.near_tile == 16) &&
                    (t.direct_tile == 2 || t.direct_tile == 4 || t.direct_tile == 8);
    if (!ok) {
        c->set_error("tuning: walk_kernel 0|1, near_tile 4|8|16, direct_tile 2|4|8");
        return NBG_E_INVALID;
    }
    return NBG_OK;
}
}  // namespace

int nbg_tuning_get(nbg_ctx* c, nbg_tuning* t) {
    if (!c || !t) return NBG_E_INVALID;
    *t = c->tune;
    return NBG_OK;
}

int nbg_tuning_set(nbg_ctx* c, const nbg_tuning* t) {
    if (!c || !t) return NBG_E_INVALID;
    if (int r = check_tuning(c, *t)) return r;
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    drop_graph(c);  // the captured step holds the old kernels
    c->tune = *t;
    c->tune.reserved = 0;
    return NBG_OK;
}

int nbg_autotune(nbg_ctx* c, const nbg_params* p, int32_t reps, nbg_tuning* best, double* ms, int32_t cap,
                 int32_t* n_out) {
    if (!c) return NBG_E_INVALID;
    if (int r = validate(c, p)) return r;
    if (reps < 1 || reps > 1000) {
        c->set_error("reps must be in 1..1000");
        return NBG_E_INVALID;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    std::vector<nbg_tuning> cand;
    const nbg_tuning cur = c->tune;
    if (p->solver == NBG_SOLVER_DIRECT) {
        for (int u : {2, 4, 8}) cand.push_back(nbg_tuning{cur.walk_kernel, cur.near_tile, u, 0});
    } else {
        for (int t : {4, 8, 16}) cand.push_back(nbg_tuning{0, t, cur.direct_tile, 0});
        // the leaf-group walk runs Plummer softening with the grouped geometric walk only
        if (!(p->flags & (NBG_FLAG_LEARNED | NBG_FLAG_PER_PARTICLE)) && p->softening == NBG_SOFT_PLUMMER &&
            p->mac == NBG_MAC_GEOMETRIC)
            cand.push_back(nbg_tuning{1, 8, cur.direct_tile, 0});
    }
    nbg_params q = *p;
    q.flags &= ~NBG_FLAG_COUNT;
    drop_graph(c);
    double best_ms = 1e300;
    nbg_tuning pick = cur;
    for (size_t k = 0; k < cand.size(); ++k) {
        c->tune = cand[k];
        if (c->n == 0) break;
        if (int r = accel_eager(c, q, nullptr)) {
            c->tune = cur;
            return r;
        }
        double sum = 0.0;
        for (int r = 0; r < reps; ++r) {
            nbg_stats st;
            if (int e = accel_eager(c, q, &st)) {
                c->tune = cur;
                return e;
            }
            sum += st.ms_walk;  // the kernel the tuning changes
        }
        const double mean = sum / reps;
        if (ms && static_cast<int32_t>(k) < cap) ms[k] = mean;
        if (mean < best_ms) {
            best_ms = mean;
            pick = cand[k];
        }
    }
    c->tune = pick;
    if (best) *best = pick;
    if (n_out) *n_out = static_cast<int32_t>(cand.size());
    return NBG_OK;
}

int nbg_diagnostics_system(nbg_ctx* c, const nbg_params* p, int32_t sys, int32_t with_potential, double* out) {
    if (!c || !out) return NBG_E_INVALID;
    if (int r = validate(c, p)) return r;
    if (sys < 0 || sys >= c->nsys) {
        c->set_error("system index out of range");
        return NBG_E_INVALID;
    }
    NBG_HIP_CHECK(c, hipSetDevice(c->device));
    int r = diagnostics(c, *p, with_potential, out, sys);
    if (r == NBG_OK) r = check_launch(c);
    return r;
}

}  // extern "C"
