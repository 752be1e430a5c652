// nbg_walk_pp.hip -- per-particle Barnes-Hut walk for gfx950.
//
// Semantics of the reference's per-particle walk WalkBH (projects/nbody_core/src/
// Solvers.cpp:181-227), used by ComputeAccelBarnesHutPerParticle (Solvers.cpp:359-382:
// every particle, geometric or relative MAC) and ComputeAccelBarnesHutTargets
// (Solvers.cpp:384-405: a target subset, geometric MAC; the force pass of block steps).
// For target i, every node reached from the root is:
//   massless -> skipped;
//   a leaf -> all its particles except i, P2P (never approximated);
//   an internal node with COM distance d and size s = 2 halfSize ->
//     relative MAC with aOld_i > 0:  accepted iff  M s^2 < alpha aOld_i d^4 / G,
//     otherwise (geometric):         accepted iff  s^2 < theta^2 d^2,
//   accepted -> monopole + quadrupole M2P, else opened (children in octant order).
// A node may be accepted by a target that lies inside it; the reference allows that too.
//
// MI355X mapping: one wavefront per 64 targets (consecutive sorted particles, or 64
// consecutive entries of a sorted-position target list), one target per lane.  The tree
// is walked once for the 64 targets: a stack entry carries the mask of targets still
// descending through it.  Up to 8 popped nodes per step are read with scalar loads and
// each lane decides the MAC for its own target; a ballot gives the accepting and the
// opening lanes.  Accepted cells and reached leaves are staged in per-wave LDS rings
// with their target masks and evaluated 64 at a time (one source per scalar load, one
// target per lane).  Decisions within FP32 rounding of the boundary are re-made in FP64
// exactly as the reference writes them, so every target's interaction list is the
// reference's.
#include <type_traits>

#include "nbg_common.hpp"
#include "nbg_internal.hpp"
#include "nbg_pair.hpp"

namespace nbg {

constexpr int kPPWaves = 4;
constexpr int kPPStack = 768;
constexpr int kPPRing = 128;  // power of two, >= 64 + kPPBatch
constexpr int kPPBatch = 8;   // nodes classified per traversal step
constexpr int kPPTile = 16;   // near-field sources loaded per batch (VGPR-resident)

struct PPArgs {
    const int32_t* tlist;   // sorted positions of the targets (null: every particle)
    const uint32_t* tcount; // device count of tlist entries
    const int32_t* bin_start;
    const int32_t* bin_count;
    int n_bins;
    int n_pad;              // index of the first zero pad entry after the sorted particles
    const int32_t* nd_start;
    const int32_t* nd_child;
    const float4* nd_rec;  // 4 per node: com, quad, (Qyz, Qzz, first particle), aux
    const double4* nd_com64;
    const double* nd_geom;
    const int32_t* sys_root;
    const int32_t* pos_sys;
    const float4* spos;
    const uint32_t* perm;
    const double* aold;     // previous |a| per particle (original index), relative MAC
    DevScalars* sc;
    float* ax;
    float* ay;
    float* az;
    float* out3;            // with tlist: results by list slot instead of ax/ay/az
    Soft soft;
    float inv_theta;
    double theta2_64;
    float alpha_f;
    int rel_fast;           // G > 0: the FP32 relative test is meaningful
    double alpha;
    double G;
    int32_t* defer_list;    // relative MAC: targets deferred to k_walk_one (count in sc->tcount)
    uint32_t* bin_cost;     // relative MAC: traversal steps of each bin in the last walk
};

// A relative-MAC target whose previous |a| is below this fraction of the largest in its
// wave is walked on its own (k_walk_one).
constexpr float kDeferRatio = 0.125f;
// A bin whose last walk took more than this many times the mean number of traversal steps
// is walked target by target (k_walk_one): the relative MAC makes every target of a bin
// near a cluster centre open deep, and one such bin would otherwise outlast the kernel.
constexpr uint32_t kHeavyBin = 3;

// Exclusive prefix over the wave of counts in 0..8, by bit-sliced ballots.
__device__ __forceinline__ int wave_excl_small8(int v, int* total) {
    int off = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint64_t m = ballot((v >> b) & 1);
        off += mbcnt(m) << b;
        tot += __popcll(m) << b;
    }
    *total = tot;
    return off;
}

// The per-particle MAC of node `nd` (com + G*mass cm, aux au) for the target at tp: the
// relative test when the target has a previous |a| (rel_lane), else the geometric test;
// FP32 first, re-made in FP64 exactly as WalkBH writes it (Solvers.cpp:204-214) when the
// FP32 result is within its rounding bound of the boundary and `exact` is set.
__device__ __forceinline__ bool pp_accept(const PPArgs& A, int nd, float4 cm, float4 au, float4 tp, float tscale,
                                          bool rel_lane, double aold, float aold_f, bool exact) {
    const float dx = cm.x - tp.x, dy = cm.y - tp.y, dz = cm.z - tp.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float cscale = fabsf(cm.x) + fabsf(cm.y) + fabsf(cm.z);
    bool accept, amb;
    if (rel_lane) {
        // G M s^2 < alpha aOld d^4  (the reference's test times G)
        const float lhs = cm.w * au.y;
        const float rhs = A.alpha_f * aold_f * (d2 * d2);
        accept = lhs < rhs;
        const float rel = 1e-5f * (cscale + tscale) * __builtin_amdgcn_rsqf(d2);
        const float window = 1e-4f * (lhs + rhs) + rhs * rel;
        amb = !(fabsf(lhs - rhs) > window) || !A.rel_fast;
    } else {
        // s^2 < theta^2 d^2  <=>  d > s / theta (the hardware square root is within 1 ulp,
        // far inside the slack below, past which the FP64 test decides)
        const float sth = __builtin_amdgcn_sqrtf(au.y) * A.inv_theta;
        const float q2 = sth * sth;
        accept = d2 > q2;
        const float slack = 2.f * sth * (1e-4f * sth + 2e-6f * (cscale + tscale));
        amb = !(fabsf(d2 - q2) > slack);
    }
    if (exact && amb) {
        const double4 c64 = A.nd_com64[nd];
        const double ex = __dsub_rn(c64.x, static_cast<double>(tp.x));
        const double ey = __dsub_rn(c64.y, static_cast<double>(tp.y));
        const double ez = __dsub_rn(c64.z, static_cast<double>(tp.z));
        const double dist2 = __dadd_rn(__dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey)), __dmul_rn(ez, ez));
        const double size = 2.0 * A.nd_geom[4 * nd + 3];
        if (rel_lane) {
            const double lhs = __dmul_rn(__dmul_rn(c64.w, size), size);
            const double rhs = __dmul_rn(__dmul_rn(__dmul_rn(A.alpha, aold), dist2), dist2) / A.G;
            accept = lhs < rhs;
        } else {
            accept = __dmul_rn(size, size) < __dmul_rn(A.theta2_64, dist2);
        }
    }
    return accept;
}

template <bool QUAD, bool KAHAN, bool SPL, bool REL, bool COUNT>
__global__ __launch_bounds__(64 * kPPWaves) void k_walk_pp(PPArgs A) {
    __shared__ int s_snode[kPPWaves][kPPStack];
    __shared__ uint64_t s_smask[kPPWaves][kPPStack];
    __shared__ __attribute__((aligned(16))) int s_fnode[kPPWaves][kPPRing];
    __shared__ __attribute__((aligned(16))) uint64_t s_fmask[kPPWaves][kPPRing];
    __shared__ int2 s_qsc[kPPWaves][kPPRing];  // near leaf: first sorted particle, count
    __shared__ uint64_t s_qmask[kPPWaves][kPPRing];
    __shared__ int s_next[kPPWaves];

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = lane_id();
    int* snode = s_snode[wave];
    uint64_t* smask = s_smask[wave];
    int* fnode = s_fnode[wave];
    uint64_t* fmask = s_fmask[wave];
    int2* qsc = s_qsc[wave];
    uint64_t* qmask = s_qmask[wave];
    const Soft soft = A.soft;
    const int n_items = A.tlist ? static_cast<int>(__builtin_amdgcn_readfirstlane(static_cast<int>(*A.tcount))) : 0;
    const int n_bins = A.tlist ? (n_items + 63) >> 6 : A.n_bins;
    // XCD-aware work distribution, as in the grouped bin walk: XCD x drains the
    // contiguous range [x per, (x+1) per) first, then helps the others
    const int xcd = blockIdx.x & 7;
    const int per = (n_bins + 7) >> 3;
    int victim = 0;
    // stack headroom: a depth-first descent pushes at most 7 net entries per level
    const int reserve = 8 * (A.sc->max_depth + 1);
    const uint64_t mybit = 1ull << lane;
    // heavy-bin threshold from the previous relative-MAC walk (0: no history)
    const uint64_t prev_cost = (REL && A.bin_cost) ? A.sc->pp_cost_prev : 0ull;
    const uint32_t heavy_steps =
        prev_cost ? static_cast<uint32_t>(min<uint64_t>(kHeavyBin * prev_cost / max(n_bins, 1), 0xffffffffull)) : 0u;

    for (;;) {
        int b = -1;
        while (victim < 8) {
            const int r = (xcd + victim) & 7;
            if (lane == 0) s_next[wave] = static_cast<int>(atomicAdd(&A.sc->work[r], 1u));
            wave_sync();
            const int k = __builtin_amdgcn_readfirstlane(s_next[wave]);
            const int lo = r * per;
            if (lo + k < min(n_bins, lo + per)) {
                b = lo + k;
                break;
            }
            ++victim;
        }
        if (b < 0) break;
        int cnt, tpos;
        if (A.tlist) {
            cnt = min(64, n_items - 64 * b);
            tpos = A.tlist[64 * b + ((lane < cnt) ? lane : 0)];
        } else {
            cnt = cld(A.bin_count, b);
            tpos = cld(A.bin_start, b) + ((lane < cnt) ? lane : 0);
        }
        const bool tv = lane < cnt;
        const float4 tp = A.spos[tpos];
        const uint32_t orig = A.perm[tpos];
        const float tscale = fabsf(tp.x) + fabsf(tp.y) + fabsf(tp.z);
        double aold = 0.0;
        if (REL) aold = tv ? A.aold[orig] : 0.0;
        const bool rel_lane = REL && aold > 0.0;
        const float aold_f = static_cast<float>(aold);
        // relative MAC: a target whose previous |a| is far below its neighbours' opens far
        // deeper than they do (near the centre of a cluster |a| -> 0 and its walk tends to a
        // direct sum); walked here it would hold the whole wave with one active lane, so it
        // is deferred to k_walk_one, which walks one target with all 64 lanes
        bool defer = false;
        const uint32_t last_cost = (REL && A.bin_cost && heavy_steps) ? cld_u32(A.bin_cost, b) : 0u;
        const bool heavy = last_cost > heavy_steps && heavy_steps > 0;
        if (REL && A.defer_list) {
            float amax = tv ? aold_f : 0.f;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
            defer = tv && (heavy || aold_f < kDeferRatio * amax);
            const uint64_t dm = ballot(defer);
            if (dm) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&A.sc->tcount, static_cast<uint32_t>(__popcll(dm)));
                base = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(base)));
                if (defer) A.defer_list[base + static_cast<uint32_t>(mbcnt(dm))] = tpos;
            }
        }
        // one stack entry per distinct system root among the wave's targets
        int sp = 0;
        {
            const int myroot = A.sys_root[A.pos_sys[tpos]];
            uint64_t rem = ballot(tv && !defer);
            while (rem) {
                const int l = __ffsll(static_cast<long long>(rem)) - 1;
                const int r = rdlane(myroot, l);
                const uint64_t m = ballot(tv && !defer && myroot == r);
                if (lane == 0) {
                    snode[sp] = r;
                    smask[sp] = m;
                }
                ++sp;
                rem &= ~m;
            }
        }
        wave_sync();

        float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
        int f_head = 0, f_tail = 0, q_head = 0, q_tail = 0;
        uint32_t n_m2p = 0, n_p2p = 0;  // this lane's interactions (COUNT)
        uint32_t steps = 0;             // traversal steps of this bin (relative MAC cost record)

        // evaluates `n` staged far cells from ring position `h`, four scalar loads ahead
        // (the ring is consumed from multiples of 64, so a group of four slots is one aligned
        // 16-byte node read and two 16-byte mask reads; slots past n repeat the first node)
        auto eval_far = [&](int h, int n) __attribute__((always_inline)) {
            float sx = 0.f, sy = 0.f, sz = 0.f;
            const int hb = h & (kPPRing - 1);
            for (int e0 = 0; e0 < n; e0 += 4) {
                int nd[4];
                bool um[4];
                float4 cm[4], q[4];
                float2 q2[4];
                const int4 n4 = *reinterpret_cast<const int4*>(fnode + hb + e0);
                const ulonglong2 m01 = *reinterpret_cast<const ulonglong2*>(fmask + hb + e0);
                const ulonglong2 m23 = *reinterpret_cast<const ulonglong2*>(fmask + hb + e0 + 2);
                const int nv[4] = {n4.x, n4.y, n4.z, n4.w};
                const uint64_t mv[4] = {m01.x, m01.y, m23.x, m23.y};
                const int nd0 = __builtin_amdgcn_readfirstlane(nv[0]);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    nd[u] = (u == 0 || e0 + u < n) ? __builtin_amdgcn_readfirstlane(nv[u]) : nd0;
                    um[u] = e0 + u < n && (mv[u] & mybit);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cm[u] = cld(A.nd_rec, 4 * nd[u]);
                    q[u] = cld(A.nd_rec, 4 * nd[u] + 1);
                    {
                        const float4 w = cld(A.nd_rec, 4 * nd[u] + 2);
                        q2[u] = make_float2(w.x, w.y);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    m2p_add<QUAD, SPL>(cm[u], q[u], q2[u], tp.x, tp.y, tp.z, soft, um[u], sx, sy, sz);
                    if (COUNT) n_m2p += um[u];
                }
            }
            kadd<KAHAN>(ax, cx, sx);
            kadd<KAHAN>(ay, cy, sy);
            kadd<KAHAN>(az, cz, sz);
        };
        // evaluates the particles of `n` staged leaves from ring position `h`, kPPTile per
        // batch across leaf boundaries; the last batch's unused slots read the zero pad
        // FAST: every pair term is finite (Plummer with eps > 0, or the spline), so a lane
        // weighs a source by its use (1 or 0) instead of selecting, and its own particle
        // adds an exact zero (d = 0); otherwise unused sources and the target itself are
        // dropped by a select.
        auto eval_near_t = [&](int h, int n, auto fast) __attribute__((always_inline)) {
            constexpr bool FAST = decltype(fast)::value;
            float sx = 0.f, sy = 0.f, sz = 0.f;
            int e = 0, k = 0, st = 0, c = 0;
            uint64_t em = 0;
            float wuse = 0.f;
            const __amdgpu_buffer_rsrc_t srec = rec_rsrc(A.spos, A.n_pad + 1);
            auto load_entry = [&](int ee) __attribute__((always_inline)) {
                const int slot = (h + ee) & (kPPRing - 1);
                const int2 e2 = qsc[slot];
                st = __builtin_amdgcn_readfirstlane(e2.x);
                c = __builtin_amdgcn_readfirstlane(e2.y);
                em = qmask[slot];
                wuse = (em & mybit) ? 1.f : 0.f;
            };
            if (n > 0) load_entry(0);
            int groups = 0;
            while (e < n) {
                int idx[kPPTile];
                float us[kPPTile];
#pragma unroll
                for (int u = 0; u < kPPTile; ++u) {
                    if (e < n) {
                        idx[u] = st + k;
                        us[u] = wuse;
                        if (COUNT) n_p2p += (em & mybit) && (st + k != tpos);
                        if (++k == c) {
                            k = 0;
                            if (++e < n) load_entry(e);
                        }
                    } else {
                        idx[u] = A.n_pad;
                        us[u] = 0.f;
                    }
                }
                float4 spt[kPPTile];
#pragma unroll
                for (int u = 0; u < kPPTile; ++u) spt[u] = bld(srec, idx[u]);
#pragma unroll
                for (int u = 0; u < kPPTile; ++u) {
                    const float4 q = spt[u];
                    const float dx = q.x - tp.x, dy = q.y - tp.y, dz = q.z - tp.z;
                    float f;
                    if (FAST) {
                        f = pair_coef_d<SPL>(dx, dy, dz, q.w * us[u], soft);
                    } else {
                        f = pair_coef<SPL>(dx * dx + dy * dy + dz * dz, q.w, soft);
                        f = (us[u] > 0.f && idx[u] != tpos) ? f : 0.f;
                    }
                    sx += f * dx;
                    sy += f * dy;
                    sz += f * dz;
                }
                // compensated flush every 64 sources
                if (++groups == 64 / kPPTile) {
                    groups = 0;
                    kadd<KAHAN>(ax, cx, sx);
                    kadd<KAHAN>(ay, cy, sy);
                    kadd<KAHAN>(az, cz, sz);
                    sx = sy = sz = 0.f;
                }
            }
            kadd<KAHAN>(ax, cx, sx);
            kadd<KAHAN>(ay, cy, sy);
            kadd<KAHAN>(az, cz, sz);
        };
        const bool near_fast = SPL || soft.eps2 > 0.f;
        auto eval_near = [&](int h, int n) __attribute__((always_inline)) {
            if (near_fast)
                eval_near_t(h, n, std::true_type{});
            else
                eval_near_t(h, n, std::false_type{});
        };

        while (sp > 0) {
            if (REL) ++steps;
            const int room = kPPStack - sp - reserve;
            const int k = min(min(kPPBatch, sp), max(1, room / 8));
            sp -= k;
            int nd[kPPBatch];
            uint64_t act[kPPBatch];
#pragma unroll
            for (int u = 0; u < kPPBatch; ++u) {
                const int e = sp + ((u < k) ? u : 0);
                nd[u] = __builtin_amdgcn_readfirstlane(snode[e]);
                const uint64_t m = smask[e];
                const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(m))));
                const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(m >> 32))));
                act[u] = (u < k) ? ((static_cast<uint64_t>(hi) << 32) | lo) : 0ull;
            }
            float4 cmv[kPPBatch], auv[kPPBatch];
#pragma unroll
            for (int u = 0; u < kPPBatch; ++u) {
                cmv[u] = cld(A.nd_rec, 4 * nd[u]);
                auv[u] = cld(A.nd_rec, 4 * nd[u] + 3);
            }
            wave_sync();  // the popped entries are read before any push overwrites them
#pragma unroll
            for (int u = 0; u < kPPBatch; ++u) {
                const uint64_t am_in = act[u];
                const float4 cm = cmv[u];
                const float4 au = auv[u];
                const uint32_t info = __float_as_uint(au.w);
                if (am_in == 0ull || !(cm.w > 0.f)) continue;  // unused slot or massless node
                if (info & kLeafBit) {
                    // a leaf: P2P for every active target (self excluded when evaluated)
                    if (lane == 0) {
                        const int slot = q_tail & (kPPRing - 1);
                        qsc[slot] = make_int2(cld(A.nd_start, nd[u]), static_cast<int>(info & ~kLeafBit));
                        qmask[slot] = am_in;
                    }
                    ++q_tail;
                    continue;
                }
                const bool mine = (am_in & mybit) != 0ull;
                const bool accept = pp_accept(A, nd[u], cm, au, tp, tscale, rel_lane, aold, aold_f, mine);
                const uint64_t am = ballot(mine && accept);
                const uint64_t om = am_in & ~am;
                if (am) {
                    if (lane == 0) {
                        const int slot = f_tail & (kPPRing - 1);
                        fnode[slot] = nd[u];
                        fmask[slot] = am;
                    }
                    ++f_tail;
                }
                if (om) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int kid = cld(A.nd_child, 8 * nd[u] + c);
                        if (kid >= 0) {
                            if (lane == 0) {
                                snode[sp] = kid;
                                smask[sp] = om;
                            }
                            ++sp;
                        }
                    }
                }
            }
            wave_sync();
            while (f_tail - f_head >= 64) {
                eval_far(f_head, 64);
                f_head += 64;
            }
            while (q_tail - q_head >= 64) {
                eval_near(q_head, 64);
                q_head += 64;
            }
        }
        eval_far(f_head, f_tail - f_head);
        eval_near(q_head, q_tail - q_head);
        if (REL && A.bin_cost && lane == 0) {
            // a deferred heavy bin keeps its recorded cost
            const uint32_t cost = heavy ? last_cost : steps;
            if (!heavy) A.bin_cost[b] = steps;
            atomicAdd(reinterpret_cast<unsigned long long*>(&A.sc->pp_cost_sum), static_cast<unsigned long long>(cost));
        }
        if (COUNT) {
            // wave totals, one atomic per wave
            uint32_t fm = tv ? n_m2p : 0u, fq = tv ? n_p2p : 0u;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                fm += static_cast<uint32_t>(__shfl_xor(static_cast<int>(fm), o, 64));
                fq += static_cast<uint32_t>(__shfl_xor(static_cast<int>(fq), o, 64));
            }
            if (lane == 0) {
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.sc->pp_m2p), static_cast<unsigned long long>(fm));
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.sc->pp_p2p), static_cast<unsigned long long>(fq));
            }
        }
        if (tv && !defer) {
            const float rx = KAHAN ? ax - cx : ax, ry = KAHAN ? ay - cy : ay, rz = KAHAN ? az - cz : az;
            if (A.out3) {
                const int slot = 64 * b + lane;
                A.out3[3 * slot + 0] = rx;
                A.out3[3 * slot + 1] = ry;
                A.out3[3 * slot + 2] = rz;
            } else {
                A.ax[orig] = rx;
                A.ay[orig] = ry;
                A.az[orig] = rz;
            }
        }
        wave_sync();
    }
}

// One target per wave, the 64 lanes walking its tree together: up to 64 stacked nodes are
// classified per step, one per lane, each lane summing the M2P and P2P terms of the nodes
// it takes; a wave reduction gives the target's acceleration.  Used for the relative-MAC
// targets deferred by k_walk_pp (their walks are far longer than their neighbours').
constexpr int kOneWaves = 4;
constexpr int kOneStack = 4096;

template <bool QUAD, bool KAHAN, bool SPL, bool COUNT>
__global__ __launch_bounds__(64 * kOneWaves) void k_walk_one(PPArgs A) {
    __shared__ int s_stack[kOneWaves][kOneStack];
    __shared__ int s_next[kOneWaves];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = lane_id();
    int* stack = s_stack[wave];
    const Soft soft = A.soft;
    const int n_items = static_cast<int>(__builtin_amdgcn_readfirstlane(static_cast<int>(A.sc->tcount)));
    const int reserve = 8 * (A.sc->max_depth + 1);
    for (;;) {
        if (lane == 0) s_next[wave] = static_cast<int>(atomicAdd(&A.sc->work2, 1u));
        wave_sync();
        const int k = __builtin_amdgcn_readfirstlane(s_next[wave]);
        if (k >= n_items) break;
        const int tpos = cld(A.defer_list, k);
        const float4 tp = cld(A.spos, tpos);
        const uint32_t orig = cld_u32(A.perm, tpos);
        const double aold = A.aold[orig];
        const bool rel_lane = aold > 0.0;
        const float aold_f = static_cast<float>(aold);
        const float tscale = fabsf(tp.x) + fabsf(tp.y) + fabsf(tp.z);
        if (lane == 0) stack[0] = cld(A.sys_root, cld(A.pos_sys, tpos));
        int sp = 1;
        wave_sync();
        float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
        uint32_t n_m2p = 0, n_p2p = 0;
        while (sp > 0) {
            const int room = kOneStack - sp - reserve;
            const int kk = min(min(64, sp), max(1, room / 7));
            const int node = (lane < kk) ? stack[sp - kk + lane] : -1;
            sp -= kk;
            wave_sync();
            int kids[8];
            int nk = 0;
            if (node >= 0) {
                const float4 cm = A.nd_rec[4 * node];
                const float4 au = A.nd_rec[4 * node + 3];
                const uint32_t info = __float_as_uint(au.w);
                if (cm.w > 0.f) {
                    if (info & kLeafBit) {
                        // a leaf: its particles except the target, P2P
                        const int st = A.nd_start[node];
                        const int cnt = static_cast<int>(info & ~kLeafBit);
                        for (int q = 0; q < cnt; ++q) {
                            const float4 sp4 = A.spos[st + q];
                            const float dx = sp4.x - tp.x, dy = sp4.y - tp.y, dz = sp4.z - tp.z;
                            float f = pair_coef<SPL>(dx * dx + dy * dy + dz * dz, sp4.w, soft);
                            f = (st + q == tpos) ? 0.f : f;
                            kadd<KAHAN>(ax, cx, f * dx);
                            kadd<KAHAN>(ay, cy, f * dy);
                            kadd<KAHAN>(az, cz, f * dz);
                            if (COUNT) n_p2p += (st + q != tpos);
                        }
                    } else if (pp_accept(A, node, cm, au, tp, tscale, rel_lane, aold, aold_f, true)) {
                        float ex, ey, ez;
                        m2p_cell<QUAD, SPL>(cm, A.nd_rec[4 * node + 1], A.nd_rec[4 * node + 3].x, tp.x, tp.y, tp.z, soft, ex, ey,
                                            ez);
                        kadd<KAHAN>(ax, cx, ex);
                        kadd<KAHAN>(ay, cy, ey);
                        kadd<KAHAN>(az, cz, ez);
                        if (COUNT) ++n_m2p;
                    } else {
                        const int4 c0 = reinterpret_cast<const int4*>(A.nd_child)[2 * node];
                        const int4 c1 = reinterpret_cast<const int4*>(A.nd_child)[2 * node + 1];
                        kids[0] = c0.x; kids[1] = c0.y; kids[2] = c0.z; kids[3] = c0.w;
                        kids[4] = c1.x; kids[5] = c1.y; kids[6] = c1.z; kids[7] = c1.w;
#pragma unroll
                        for (int c = 0; c < 8; ++c) nk += kids[c] >= 0;
                    }
                }
            }
            int ktot;
            const int koff = wave_excl_small8(nk, &ktot);
            if (nk > 0) {
                int w = sp + koff;
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (kids[c] >= 0) stack[w++] = kids[c];
            }
            sp += ktot;
            wave_sync();
        }
        const float rx = wave_sum(KAHAN ? ax - cx : ax);
        const float ry = wave_sum(KAHAN ? ay - cy : ay);
        const float rz = wave_sum(KAHAN ? az - cz : az);
        if (COUNT) {
            uint32_t fm = n_m2p, fq = n_p2p;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                fm += static_cast<uint32_t>(__shfl_xor(static_cast<int>(fm), o, 64));
                fq += static_cast<uint32_t>(__shfl_xor(static_cast<int>(fq), o, 64));
            }
            if (lane == 0) {
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.sc->pp_m2p), static_cast<unsigned long long>(fm));
                atomicAdd(reinterpret_cast<unsigned long long*>(&A.sc->pp_p2p), static_cast<unsigned long long>(fq));
            }
        }
        if (lane == 0) {
            A.ax[orig] = rx;
            A.ay[orig] = ry;
            A.az[orig] = rz;
        }
        wave_sync();
    }
}

// |a| per particle in FP64 from the current accelerations: the previous-step magnitude the
// relative MAC reads (System::CaptureAOld, projects/nbody_core/include/ngrav/System.hpp:302-310).
__global__ __launch_bounds__(256) void k_aold(const float* __restrict__ ax, const float* __restrict__ ay,
                                              const float* __restrict__ az, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = ax[i], y = ay[i], z = az[i];
    out[i] = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z)));
}

void launch_aold(nbg_ctx* c) {
    hipLaunchKernelGGL(k_aold, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->ax, c->ay, c->az, c->n,
                       c->aold);
}

template <bool SPL>
void dispatch_walk_one(nbg_ctx* c, const PPArgs& A, bool quad, bool kahan, bool count) {
    const dim3 grid(8 * 32 * 2), block(64 * kOneWaves);
    if (quad) {
        if (count) hipLaunchKernelGGL((k_walk_one<true, true, SPL, true>), grid, block, 0, c->stream, A);
        else if (kahan) hipLaunchKernelGGL((k_walk_one<true, true, SPL, false>), grid, block, 0, c->stream, A);
        else hipLaunchKernelGGL((k_walk_one<true, false, SPL, false>), grid, block, 0, c->stream, A);
    } else {
        if (count) hipLaunchKernelGGL((k_walk_one<false, true, SPL, true>), grid, block, 0, c->stream, A);
        else if (kahan) hipLaunchKernelGGL((k_walk_one<false, true, SPL, false>), grid, block, 0, c->stream, A);
        else hipLaunchKernelGGL((k_walk_one<false, false, SPL, false>), grid, block, 0, c->stream, A);
    }
}

template <bool SPL, bool REL>
void dispatch_walk_pp(nbg_ctx* c, const PPArgs& A, bool quad, bool kahan, bool count) {
    // 8 XCDs x 32 CUs x 3 resident blocks (51 KB of LDS each); interaction counting is
    // built for the compensated variants
    const dim3 grid(8 * 32 * 3), block(64 * kPPWaves);
    if (quad) {
        if (count) hipLaunchKernelGGL((k_walk_pp<true, true, SPL, REL, true>), grid, block, 0, c->stream, A);
        else if (kahan) hipLaunchKernelGGL((k_walk_pp<true, true, SPL, REL, false>), grid, block, 0, c->stream, A);
        else hipLaunchKernelGGL((k_walk_pp<true, false, SPL, REL, false>), grid, block, 0, c->stream, A);
    } else {
        if (count) hipLaunchKernelGGL((k_walk_pp<false, true, SPL, REL, true>), grid, block, 0, c->stream, A);
        else if (kahan) hipLaunchKernelGGL((k_walk_pp<false, true, SPL, REL, false>), grid, block, 0, c->stream, A);
        else hipLaunchKernelGGL((k_walk_pp<false, false, SPL, REL, false>), grid, block, 0, c->stream, A);
    }
}

void launch_walk_pp(nbg_ctx* c, const nbg_params& p, const int32_t* tlist, const uint32_t* tcount, float* out3) {
    PPArgs A;
    A.tlist = tlist;
    A.tcount = tcount;
    A.bin_start = c->bin_start;
    A.bin_count = c->bin_count;
    A.n_bins = c->n_bins;
    A.n_pad = c->n;
    A.nd_start = c->nd_start;
    A.nd_child = c->nd_child;
    A.nd_rec = c->nd_rec;
    A.nd_com64 = c->nd_com64;
    A.nd_geom = c->nd_geom;
    A.sys_root = c->sys_root;
    A.pos_sys = c->pos_sys;
    A.spos = c->spos;
    A.perm = c->perm;
    A.aold = c->aold;
    A.sc = c->dsc;
    A.ax = c->ax;
    A.ay = c->ay;
    A.az = c->az;
    A.out3 = out3;
    A.soft = make_soft(p);
    A.inv_theta = (p.theta > 0.0) ? static_cast<float>(1.0 / p.theta) : __builtin_inff();
    A.theta2_64 = p.theta * p.theta;
    A.alpha = p.alpha;
    A.alpha_f = static_cast<float>(p.alpha);
    A.G = p.G;
    A.rel_fast = p.G > 0.0 ? 1 : 0;
    const bool quad = p.multipole == NBG_QUADRUPOLE;
    const bool kahan = (p.flags & NBG_FLAG_KAHAN) != 0;
    // target lists (block steps, nbg_accel_bh_targets) always take the geometric MAC
    const bool rel = p.mac == NBG_MAC_RELATIVE && tlist == nullptr;
    A.defer_list = nullptr;
    A.bin_cost = nullptr;
    const bool spl = p.softening == NBG_SOFT_SPLINE;
    // NBG_FLAG_COUNT counts with compensated sums (the counts do not depend on the sums)
    const bool count = (p.flags & NBG_FLAG_COUNT) != 0;
    // relative MAC over every particle: outlier targets go to the list (its count, zeroed
    // by the build, in sc->tcount) and are walked one per wave afterwards
    A.defer_list = rel ? c->tlist : nullptr;
    A.bin_cost = rel ? c->bin_cost : nullptr;
    if (spl) {
        if (rel) dispatch_walk_pp<true, true>(c, A, quad, kahan, count);
        else dispatch_walk_pp<true, false>(c, A, quad, kahan, count);
        if (rel) dispatch_walk_one<true>(c, A, quad, kahan, count);
    } else {
        if (rel) dispatch_walk_pp<false, true>(c, A, quad, kahan, count);
        else dispatch_walk_pp<false, false>(c, A, quad, kahan, count);
        if (rel) dispatch_walk_one<false>(c, A, quad, kahan, count);
    }
}

}  // namespace nbg
