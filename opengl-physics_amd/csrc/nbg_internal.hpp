// nbg_internal.hpp -- context layout and kernel launchers shared by the .hip files of
// libnbg_hip.so.  Everything here is gfx950 (CDNA4) code: 64-wide wavefronts, LDS, HBM3E.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nbg.h"

namespace nbg {

// ---------------------------------------------------------------------------- constants
constexpr int kNcrit = 8;           // AdaptiveTree ncrit (projects/nbody_core/include/ngrav/Tree.hpp:38)
constexpr int kMaxDepth = 40;       // AdaptiveTree maxDepth (same line)
constexpr int kLevelsTiny = 13;     // levels of the first key tried: 39 bits, 5 sort passes
constexpr int kLevelsFast = 16;     // the next tier: 48 bits, 6 sort passes
constexpr int kLevelsShort = 21;    // levels held by one 63-bit key
constexpr int kLevelsDeep = 40;     // levels held by the two-word key
// Depths that hold internal nodes for keys of L levels: 0..L-1 (a depth-L cell of a short
// key is a leaf, and so is every depth-40 cell of the two-word key).
template <int L>
constexpr int kInternalDepths = (L <= kLevelsShort) ? L : kMaxDepth;
constexpr int kSortThreads = 256;   // radix-sort workgroup
constexpr int kSortItems = 8;       // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;  // keys per sort block (2048)
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

// Leaf info word stored (as float bits) in node_aux.w: bit 31 set = leaf, low bits =
// particle count (leaf) or child count (internal).
constexpr uint32_t kLeafBit = 0x80000000u;
// Blocks of the grid-stride node kernels (k_node_finish, k_depth_scatter): one scan tile of
// per-block depth counts per depth.
constexpr int kNodeBlocks = 2048;

// Zero float4 entries after the sorted particles: the walk loads 8 particles at a time
// from a leaf start without a bound check.
constexpr int kSposPad = 8;
// Largest context: sorted-particle indices (padding included) stay below 2^28, so the
// force kernels' 32-bit byte offsets (index << 4) into the sorted particles cannot wrap.
constexpr int64_t kMaxParticles = (int64_t{1} << 28) - 1024;

// Device-side scalars written by the build, read back only for stats.
struct DevScalars {
    int32_t n_nodes;
    int32_t n_leaves;
    int32_t max_depth;
    int32_t overflow;    // a depth-21 cell held more than ncrit particles
    int32_t cap_exceeded;
    int32_t sticky_overflow;  // like overflow/cap_exceeded, but only cleared by the host
    int32_t sticky_cap;
    int32_t pad;
    uint32_t depth_count[64];   // internal nodes per depth
    uint32_t depth_offset[64];  // exclusive scan of depth_count
    uint32_t depth_fill[64];    // scatter cursors
    uint32_t work[8];           // per-XCD work counters of the walk
    uint32_t learn_pairs;       // (leaf, cell) pairs listed by the walk for the learned correction
    int32_t sticky_learn;       // the pair list overflowed during a graph batch (host clears)
    int32_t learn_fix_e;        // learned correction: fixed-point sums hold v * 2^learn_fix_e
    int32_t learn_fix_over;     // a correction term exceeded the one-pass fixed-point range
    double dt_step;             // adaptive steps: dt of the current step
    double t_accum;             // adaptive steps: time advanced since the host cleared it
    uint64_t particle_evals;    // block steps: targets evaluated since the host cleared it
    uint32_t tcount;            // entries of the current target list
    int32_t rung_max;           // block steps: deepest rung of the coarse step
    uint64_t pp_m2p;            // per-particle walk interactions (NBG_FLAG_COUNT), zeroed by the build
    uint64_t pp_p2p;
    uint32_t work2;             // work counter of the one-target-per-wave walk, zeroed by the build
    uint32_t pad3;
    uint64_t pp_cost_sum;       // relative-MAC walk: traversal steps over all bins (this walk)
    uint64_t pp_cost_prev;      // the same of the previous walk (the build moves sum here)
};

// ---------------------------------------------------------------------------- helpers
#define NBG_HIP_CHECK(ctx, expr)                                                         \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            (ctx)->set_error(std::string(#expr) + ": " + hipGetErrorString(e_));         \
            return NBG_E_HIP;                                                            \
        }                                                                                \
    } while (0)

template <typename T>
inline T* dev_alloc(size_t count, hipError_t& err) {
    void* p = nullptr;
    if (count == 0) count = 1;
    err = hipMalloc(&p, count * sizeof(T));
    if (err != hipSuccess) return nullptr;
    return static_cast<T*>(p);
}

inline int grid_for(int64_t n, int threads) { return static_cast<int>((n + threads - 1) / threads); }

}  // namespace nbg

// ---------------------------------------------------------------------------- context
struct nbg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    void set_error(const std::string& s) { err = s; }

    int64_t cap_n = 0;    // particle capacity
    int64_t cap_nodes = 0;
    int n = 0;            // particles uploaded
    int nsys = 0;         // systems in the batch
    bool have_state = false;
    bool have_tree = false;
    int last_levels = nbg::kLevelsTiny;  // key levels of the last tree
    int key_levels = nbg::kLevelsTiny;   // key levels a tree build tries first (raised after an overflow)
    int graph_levels = 0;           // key levels the captured step graph was built with
    double last_G = 1.0;  // G folded into the device masses of the last evaluation

    // batch layout (host + device)
    std::vector<int64_t> h_sys_off;
    int32_t* sys_off = nullptr;     // nsys+1
    int32_t* pos_sys = nullptr;     // system of each particle / sorted position
    int32_t* sys_root = nullptr;    // root node of each system
    double* sys_geom = nullptr;     // 4 per system: root centre x,y,z, halfSize
    uint32_t* blk_bbox = nullptr;   // 6 per sort block: order-preserving min x,y,z, max x,y,z
    // sort blocks: each covers at most kSortTile keys of one system
    int n_blocks = 0;
    int32_t* blk_sys = nullptr;
    int32_t* blk_start = nullptr;
    int32_t* blk_count = nullptr;
    int32_t* sys_blk0 = nullptr;    // nsys+1: first sort block of each system
    // walk bins: runs of <= 64 consecutive sorted positions inside one system
    int n_bins = 0;
    int32_t* bin_start = nullptr;
    int32_t* bin_count = nullptr;
    uint32_t* bin_cost = nullptr;   // relative-MAC traversal steps of each bin, last walk

    // particle state, original order (FP32 SoA)
    float *x = nullptr, *y = nullptr, *z = nullptr, *m = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr;
    float *ax = nullptr, *ay = nullptr, *az = nullptr;
    // backup of the state for the deep-key re-run of a graph batch
    float* backup = nullptr;
    // FP64 positions (nbg_upload_f64, allocated on first use): the tree is built from them
    // and the grouped walk differences positions relative to leaf centres (srel)
    bool f64_state = false;
    double *x64 = nullptr, *y64 = nullptr, *z64 = nullptr;
    uint64_t* blk_bbox64 = nullptr;  // 6 per sort block, order-preserving FP64 bounds
    float4* srel = nullptr;          // per sorted position: offset from its leaf's centre, G*m (+ zero pad)

    // keys and sort
    uint64_t *key_a = nullptr, *key_b = nullptr;   // levels 1..21 (ping-pong)
    uint64_t *klo_a = nullptr, *klo_b = nullptr;   // levels 22..40 (deep path)
    uint32_t *val_a = nullptr, *val_b = nullptr;
    uint32_t* hist = nullptr;                      // kRadix * n_blocks
    uint64_t* scan_part = nullptr;                 // block partials of the scans
    uint64_t* key_sorted = nullptr;                // final sorted keys (alias of a or b)
    uint64_t* klo_sorted = nullptr;
    uint32_t* perm = nullptr;                      // sorted position -> original index

    // build
    int32_t* leaf_depth = nullptr;   // per sorted position
    int32_t* cpl_prev = nullptr;     // common-prefix length with the previous position
    uint64_t* node_scan = nullptr;   // (nodes started << 32 | leaf started), then its exclusive scan
    int32_t* blk_depth = nullptr;    // deepest leaf per k_leaf_depth block

    // node arrays (DFS pre-order ids)
    int32_t* nd_start = nullptr;
    int32_t* nd_count = nullptr;
    int32_t* nd_depth = nullptr;
    int32_t* nd_parent = nullptr;
    int32_t* nd_oct = nullptr;
    int32_t* nd_child = nullptr;     // 8 per node
    double* nd_geom = nullptr;       // 4 per node: centre x,y,z, halfSize (FP64)
    float4* nd_rec = nullptr;        // 4 per node (one 64-byte line): (com, G*M), G*(Qxx, Qyy, Qxy, Qxz),
                                     // G*(Qyz, Qzz, first sorted particle, -), aux (G*Qyz, size^2, radius, info)
    double4* nd_com64 = nullptr;     // FP64 com x,y,z and mass (without G): exact MAC decisions
    double* nd_rmax64 = nullptr;     // FP64 maxRadius of leaves (the group radius of the MAC)
    int32_t* leaf_list = nullptr;    // leaf node ids, in pre-order
    int32_t* leaf_of = nullptr;      // sorted position -> its leaf node id
    int32_t* depth_nodes = nullptr;  // internal node ids grouped by depth
    uint32_t* dep_hist = nullptr;    // internal nodes per (depth, node-kernel block), depth-major
    uint32_t* dep_off = nullptr;     // its exclusive scan: first depth_nodes slot per (depth, block)
    uint32_t* dep_part = nullptr;    // the scan's tile sums: internal nodes per depth
    int32_t* cnt_far = nullptr;      // per node (leaf groups), NBG_FLAG_COUNT
    int32_t* cnt_near = nullptr;

    // sorted particles: x,y,z, G*m
    float4* spos = nullptr;
    float4* stage4 = nullptr;        // N records: x,y,z, G*m in input order for the sorted gather,
                                     // then the learned walk's accelerations in sorted order

    // scalars
    nbg::DevScalars* dsc = nullptr;
    nbg::DevScalars* hsc = nullptr;   // pinned mirror

    // diagnostics scratch
    double* diag_part = nullptr;
    int diag_part_cap = 0;

    // per-particle walk, adaptive and block steps
    double* aold = nullptr;          // previous |a| per particle (relative MAC)
    int32_t* rung = nullptr;         // block-step rung per particle (original order)
    int32_t* tlist = nullptr;        // target list (sorted positions or original indices)
    uint32_t* list_off = nullptr;    // per-tile counts / offsets of the list compaction
    double* red_part = nullptr;      // block partials of the dt / rung reductions
    bool have_rungs = false;

    // timing
    hipEvent_t ev[14] = {};

    // graph of one step
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    nbg_params graph_params{};
    bool graph_valid = false;
    bool capturing = false;  // inside hipStreamBeginCapture: events become graph nodes
    nbg_tuning tune{0, 16, 4, 0};  // walk kernel, near P2P tile, direct P2P tile (autotuner)

    // learned far-field correction (allocated on first use)
    float4* nd_mom = nullptr;       // 4 per node: second (6) and third (10) moments about the FP32 com, G*m weights
    int learn_act = NBG_LEARN_ACT_SILU;  // activation of the installed model (blob[NBG_LEARN_ACT_SLOT])
    bool learn_two_pass = false;    // corrections summed with per-leaf fixed-point units (after a range overflow)
    long long* lacc = nullptr;      // per node, 4 words: a leaf group's correction as exact fixed-point sums
    uint32_t* lunit = nullptr;      // per node: bits of the largest |component| of its pairs' corrections
    int2* lpairs = nullptr;         // (leaf, cell) pairs listed by the walk
    float4* lpv = nullptr;          // per listed pair: its correction (x, y, z)
    int64_t lpairs_cap = 0;
    float* lw32 = nullptr;          // FP32 model blob (NBG_LEARN_NPARAM)
    uint16_t* lw16 = nullptr;       // BF16 copy of the trainable part (blob order)
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int64_t adam_t = 0;
    bool have_weights = false;
    float* pool_feat = nullptr;     // 16 per sample
    float* pool_tgt = nullptr;      // 4 per sample: residual x,y,z, weight
    int64_t pool_n = 0, pool_cap = 0;
    float pool_inv_wmean = 1.f;
    float* amag2 = nullptr;         // |a_direct|^2 per particle (sample collection)
    float* grad_part = nullptr;     // per training block partial gradients
    float* loss_part = nullptr;
    float* loss_dev = nullptr;
    int grad_blocks_cap = 0;
    int loss_cap = 0;
};

namespace nbg {

// ---------------------------------------------------------------------------- launchers
// tree.hip
void launch_bbox_and_keys(nbg_ctx* c, int levels);
void launch_sort(nbg_ctx* c, int levels);
void launch_build(nbg_ctx* c, int levels);
void launch_upsweep(nbg_ctx* c, int levels, int multipole, double G, bool moments);
void launch_fill_i32(nbg_ctx* c, int32_t* p, int64_t n, int32_t v);
// walk.hip
void launch_walk(nbg_ctx* c, const nbg_params& p);
void launch_direct(nbg_ctx* c, const nbg_params& p);
void launch_direct_targets(nbg_ctx* c, const nbg_params& p, int64_t nt, const int32_t* targets,
                           const uint32_t* nt_dev, float* out3);
void launch_pack(nbg_ctx* c, double G);
void launch_rel(nbg_ctx* c);  // srel from the FP64 positions and the built tree
void launch_direct_raw(hipStream_t stream, const nbg_params& p, int n, const float* x, const float* y,
                       const float* z, const float* m, float4* stage, float* ax, float* ay, float* az);
// walk_pp.hip
void launch_walk_pp(nbg_ctx* c, const nbg_params& p, const int32_t* tlist, const uint32_t* tcount, float* out3);
void launch_aold(nbg_ctx* c);
// steps.hip
void launch_adaptive_dt(nbg_ctx* c, const nbg_params& p);
void launch_kick_drift_dev(nbg_ctx* c);
void launch_kick_dev(nbg_ctx* c);
void launch_rungs(nbg_ctx* c, const nbg_params& p);
void launch_block_kick(nbg_ctx* c, const nbg_params& p, int sub, int rmax, bool closing);
void launch_drift(nbg_ctx* c, float dt);
void launch_active_list(nbg_ctx* c, int sub, int rmax, bool sorted);
constexpr int kReduceBlocks = 1024;  // partials of the steps' reductions
// integrate.hip
void launch_kick(nbg_ctx* c, float hdt);
void launch_kick_drift(nbg_ctx* c, float hdt, float dt);
void launch_copy_f32(nbg_ctx* c, float* dst, const float* src, int64_t n);
int diagnostics(nbg_ctx* c, const nbg_params& p, int with_potential, double* out, int sys);
int64_t diag_partials_needed(int64_t n);
// learn.hip
int learn_alloc(nbg_ctx* c);
int learn_grow_pairs(nbg_ctx* c, int64_t need);
void launch_learn_zero(nbg_ctx* c);
void launch_learn_correct(nbg_ctx* c, const nbg_params& p);
void launch_learn_apply(nbg_ctx* c);  // sorted walk accelerations + leaf correction -> input order
void launch_learn_samples(nbg_ctx* c, const nbg_params& p, int64_t npairs, int64_t nnew, int64_t dst);
void launch_learn_amag2(nbg_ctx* c);
void launch_learn_stats(nbg_ctx* c, int64_t count, double* out /* device, 16+16+1 */);
void launch_learn_fwdbwd(nbg_ctx* c, const float* feat, const float* tgt, int64_t n_src, int nblocks, uint64_t seed,
                         int step, int use_hash, float inv_wmean, float inv_n, float* out);
void launch_learn_adam(nbg_ctx* c, int nblocks, float lr, int64_t t, float* loss_slot);
void launch_learn_reduce(nbg_ctx* c, int nblocks, float* grad_out, float* loss_out);
// walk.hip: learned-mode walk parameters live in nbg_params (learn_band)

}  // namespace nbg
