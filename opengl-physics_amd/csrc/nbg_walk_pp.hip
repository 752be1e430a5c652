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
#include "nbg_common.hpp"
#include "nbg_internal.hpp"
#include "nbg_pair.hpp"

namespace nbg {

constexpr int kPPWaves = 4;
constexpr int kPPStack = 768;
constexpr int kPPRing = 128;  // power of two, >= 64 + kPPBatch
constexpr int kPPBatch = 8;   // nodes classified per traversal step

struct PPArgs {
    const int32_t* tlist;   // sorted positions of the targets (null: every particle)
    const uint32_t* tcount; // device count of tlist entries
    const int32_t* bin_start;
    const int32_t* bin_count;
    int n_bins;
    int n_pad;              // index of the first zero pad entry after the sorted particles
    const int32_t* nd_start;
    const int32_t* nd_child;
    const float4* nd_com;
    const float4* nd_quad;
    const float4* nd_aux;
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
};

template <bool QUAD, bool KAHAN, bool SPL, bool REL>
__global__ __launch_bounds__(64 * kPPWaves) void k_walk_pp(PPArgs A) {
    __shared__ int s_snode[kPPWaves][kPPStack];
    __shared__ uint64_t s_smask[kPPWaves][kPPStack];
    __shared__ int s_fnode[kPPWaves][kPPRing];
    __shared__ uint64_t s_fmask[kPPWaves][kPPRing];
    __shared__ int s_qstart[kPPWaves][kPPRing];
    __shared__ int s_qcnt[kPPWaves][kPPRing];
    __shared__ uint64_t s_qmask[kPPWaves][kPPRing];
    __shared__ int s_next[kPPWaves];

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = lane_id();
    int* snode = s_snode[wave];
    uint64_t* smask = s_smask[wave];
    int* fnode = s_fnode[wave];
    uint64_t* fmask = s_fmask[wave];
    int* qstart = s_qstart[wave];
    int* qcnt = s_qcnt[wave];
    uint64_t* qmask = s_qmask[wave];
    const Soft soft = A.soft;
    const float inv_theta = A.inv_theta;
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
        const uint64_t tmask = ballot(tv);
        const float4 tp = A.spos[tpos];
        const uint32_t orig = A.perm[tpos];
        const float tscale = fabsf(tp.x) + fabsf(tp.y) + fabsf(tp.z);
        double aold = 0.0;
        if (REL) aold = tv ? A.aold[orig] : 0.0;
        const bool rel_lane = REL && aold > 0.0;
        const float aold_f = static_cast<float>(aold);
        // one stack entry per distinct system root among the wave's targets
        int sp = 0;
        {
            const int myroot = A.sys_root[A.pos_sys[tpos]];
            uint64_t rem = tmask;
            while (rem) {
                const int l = __ffsll(static_cast<long long>(rem)) - 1;
                const int r = __builtin_amdgcn_readfirstlane(__shfl(myroot, l, 64));
                const uint64_t m = ballot(tv && myroot == r);
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

        // evaluates `n` staged far cells from ring position `h`, four scalar loads ahead
        auto eval_far = [&](int h, int n) {
            float sx = 0.f, sy = 0.f, sz = 0.f;
            for (int e0 = 0; e0 < n; e0 += 4) {
                int nd[4];
                float um[4];
                float4 cm[4], q[4];
                float qyz[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int slot = (h + min(e0 + u, n - 1)) & (kPPRing - 1);
                    nd[u] = __builtin_amdgcn_readfirstlane(fnode[slot]);
                    um[u] = (e0 + u < n && (fmask[slot] & mybit)) ? 1.f : 0.f;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cm[u] = cld(A.nd_com, nd[u]);
                    q[u] = cld(A.nd_quad, nd[u]);
                    qyz[u] = __uint_as_float(cld_u32(A.nd_aux, 4 * nd[u]));
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    float ex, ey, ez;
                    m2p_cell<QUAD, SPL>(cm[u], q[u], qyz[u], tp.x, tp.y, tp.z, soft, ex, ey, ez);
                    sx += (um[u] > 0.f) ? ex : 0.f;
                    sy += (um[u] > 0.f) ? ey : 0.f;
                    sz += (um[u] > 0.f) ? ez : 0.f;
                }
            }
            kadd<KAHAN>(ax, cx, sx);
            kadd<KAHAN>(ay, cy, sy);
            kadd<KAHAN>(az, cz, sz);
        };
        // evaluates the particles of `n` staged leaves from ring position `h`, eight per
        // batch across leaf boundaries; the last batch's unused slots read the zero pad
        auto eval_near = [&](int h, int n) {
            float sx = 0.f, sy = 0.f, sz = 0.f;
            int e = 0, k = 0, st = 0, c = 0;
            uint64_t em = 0;
            auto load_entry = [&](int ee) {
                const int slot = (h + ee) & (kPPRing - 1);
                st = __builtin_amdgcn_readfirstlane(qstart[slot]);
                c = __builtin_amdgcn_readfirstlane(qcnt[slot]);
                em = qmask[slot];
            };
            if (n > 0) load_entry(0);
            int groups = 0;
            while (e < n) {
                int idx[8];
                float us[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (e < n) {
                        idx[u] = st + k;
                        us[u] = (em & mybit) ? 1.f : 0.f;
                        if (++k == c) {
                            k = 0;
                            if (++e < n) load_entry(e);
                        }
                    } else {
                        idx[u] = A.n_pad;
                        us[u] = 0.f;
                    }
                }
                float4 spt[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) spt[u] = cld(A.spos, idx[u]);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float4 q = spt[u];
                    const float dx = q.x - tp.x, dy = q.y - tp.y, dz = q.z - tp.z;
                    float f = pair_coef<SPL>(dx * dx + dy * dy + dz * dz, q.w, soft);
                    f = (us[u] > 0.f && idx[u] != tpos) ? f : 0.f;
                    sx += f * dx;
                    sy += f * dy;
                    sz += f * dz;
                }
                // compensated flush every 64 sources
                if (++groups == 8) {
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

        while (sp > 0) {
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
                cmv[u] = cld(A.nd_com, nd[u]);
                auv[u] = cld(A.nd_aux, nd[u]);
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
                        qstart[slot] = cld(A.nd_start, nd[u]);
                        qcnt[slot] = static_cast<int>(info & ~kLeafBit);
                        qmask[slot] = am_in;
                    }
                    ++q_tail;
                    continue;
                }
                const bool mine = (am_in & mybit) != 0ull;
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
                    // s^2 < theta^2 d^2  <=>  d > s / theta
                    const float sth = sqrtf(au.y) * inv_theta;
                    const float q2 = sth * sth;
                    accept = d2 > q2;
                    const float slack = 2.f * sth * (1e-4f * sth + 2e-6f * (cscale + tscale));
                    amb = !(fabsf(d2 - q2) > slack);
                }
                if (mine && amb) {
                    // FP64 decision exactly as WalkBH writes it (Solvers.cpp:204-214)
                    const double4 c64 = A.nd_com64[nd[u]];
                    const double ex = __dsub_rn(c64.x, static_cast<double>(tp.x));
                    const double ey = __dsub_rn(c64.y, static_cast<double>(tp.y));
                    const double ez = __dsub_rn(c64.z, static_cast<double>(tp.z));
                    const double dist2 =
                        __dadd_rn(__dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey)), __dmul_rn(ez, ez));
                    const double size = 2.0 * A.nd_geom[4 * nd[u] + 3];
                    if (rel_lane) {
                        const double lhs = __dmul_rn(__dmul_rn(c64.w, size), size);
                        const double rhs = __dmul_rn(__dmul_rn(__dmul_rn(A.alpha, aold), dist2), dist2) / A.G;
                        accept = lhs < rhs;
                    } else {
                        accept = __dmul_rn(size, size) < __dmul_rn(A.theta2_64, dist2);
                    }
                }
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
        if (tv) {
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

template <bool SPL, bool REL>
void dispatch_walk_pp(nbg_ctx* c, const PPArgs& A, bool quad, bool kahan) {
    // 8 XCDs x 32 CUs x 3 resident blocks (51 KB of LDS each)
    const dim3 grid(8 * 32 * 3), block(64 * kPPWaves);
    if (quad) {
        if (kahan) hipLaunchKernelGGL((k_walk_pp<true, true, SPL, REL>), grid, block, 0, c->stream, A);
        else hipLaunchKernelGGL((k_walk_pp<true, false, SPL, REL>), grid, block, 0, c->stream, A);
    } else {
        if (kahan) hipLaunchKernelGGL((k_walk_pp<false, true, SPL, REL>), grid, block, 0, c->stream, A);
        else hipLaunchKernelGGL((k_walk_pp<false, false, SPL, REL>), grid, block, 0, c->stream, A);
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
    A.nd_com = c->nd_com;
    A.nd_quad = c->nd_quad;
    A.nd_aux = c->nd_aux;
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
    const bool spl = p.softening == NBG_SOFT_SPLINE;
    if (spl) {
        if (rel) dispatch_walk_pp<true, true>(c, A, quad, kahan); else dispatch_walk_pp<true, false>(c, A, quad, kahan);
    } else {
        if (rel) dispatch_walk_pp<false, true>(c, A, quad, kahan); else dispatch_walk_pp<false, false>(c, A, quad, kahan);
    }
}

}  // namespace nbg
