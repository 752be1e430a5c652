// nbg_walk.hip -- Barnes-Hut group walk and tiled direct sum for gfx950.
//
// Walk semantics follow the reference's grouped walk exactly:
//   CollectGroupInteractions  projects/nbody_core/src/Solvers.cpp:248-277
//   ComputeAccelBarnesHutGroup projects/nbody_core/src/Solvers.cpp:279-355
// For each leaf g (the "group"), every node reached from the root is classified:
// g itself -> skipped; massless -> skipped; a leaf -> near (all its particles, P2P,
// never approximated); an internal node with worst = |com_n - com_g| - maxRadius_g > 0
// and (2 halfSize)^2 < theta^2 worst^2 -> far (monopole + quadrupole); else opened.
// Each member of g then sums far M2P, near P2P and own-leaf P2P (self excluded).
//
// MI355X mapping: one wavefront per group.  The wave walks the tree cooperatively --
// up to 64 stacked nodes are classified per step, one per lane -- and stages the far
// cells and near particles it finds in per-wave LDS rings.  Whenever 64 sources are
// staged they are evaluated with one source per lane against all of the group's
// targets (<= 8, held in registers), so source data is loaded once per group and the
// arithmetic is lane-dense.  A wave reduction sums each target over the lanes at the end.
// The pair kernels are the reference's (Solvers.cpp:19-38 softened P2P/M2P monopole,
// Solvers.cpp:147-179 unsoftened traceless quadrupole) in FP32, with optional
// compensated (Kahan) accumulation.
#include <type_traits>

#include "nbg_common.hpp"
#include "nbg_internal.hpp"
#include "nbg_pair.hpp"

namespace nbg {

constexpr int kWalkWaves = 4;
constexpr int kStackCap = 2048;
constexpr int kFarRing = 128;     // power of two
constexpr int kNearRing = 1024;   // power of two, >= 64 + 64 * kNcrit
constexpr int kMaxT = kNcrit;     // targets per pass (a depth-capped leaf takes several passes)
// Batch pops only while a full batch of 64 opened nodes (7 x 64 net pushes) plus a
// depth-first descent of 7 entries per level still fits; otherwise pop one node.
constexpr int kBatchThresh = kStackCap - 8 * (kMaxDepth + 1) - 7 * 64 - 64;

struct Acc {
    float x[kMaxT], y[kMaxT], z[kMaxT];
    float cx[kMaxT], cy[kMaxT], cz[kMaxT];  // Kahan compensation
};

// P2P source (sx,sy,sz, G*m) onto all targets.  `self` = target slot to skip (-1 none).
template <bool KAHAN>
__device__ __forceinline__ void p2p(Acc& a, const float* tx, const float* ty, const float* tz, int nt, float4 s,
                                    float eps2, int self) {
#pragma unroll
    for (int t = 0; t < kMaxT; ++t) {
        if (t < nt) {
            const float dx = s.x - tx[t], dy = s.y - ty[t], dz = s.z - tz[t];
            const float r2 = dx * dx + dy * dy + dz * dz + eps2;
            const float inv = __builtin_amdgcn_rsqf(r2);
            float f = s.w * (inv * inv * inv);
            f = (t == self) ? 0.f : f;
            kadd<KAHAN>(a.x[t], a.cx[t], f * dx);
            kadd<KAHAN>(a.y[t], a.cy[t], f * dy);
            kadd<KAHAN>(a.z[t], a.cz[t], f * dz);
        }
    }
}

// M2P: softened monopole of G*mass at com, plus the unsoftened quadrupole of G*Q.
template <bool QUAD, bool KAHAN>
__device__ __forceinline__ void m2p(Acc& a, const float* tx, const float* ty, const float* tz, int nt, float4 cm,
                                    float4 q, float qyz, float eps2) {
#pragma unroll
    for (int t = 0; t < kMaxT; ++t) {
        if (t < nt) {
            const float dx = cm.x - tx[t], dy = cm.y - ty[t], dz = cm.z - tz[t];
            const float dist2 = dx * dx + dy * dy + dz * dz;
            const float inv = __builtin_amdgcn_rsqf(dist2 + eps2);
            const float f = cm.w * (inv * inv * inv);
            float ex = f * dx, ey = f * dy, ez = f * dz;
            if (QUAD) {
                const float qxx = q.x, qyy = q.y, qxy = q.z, qxz = q.w, qzz = -(qxx + qyy);
                const float qdx = qxx * dx + qxy * dy + qxz * dz;
                const float qdy = qxy * dx + qyy * dy + qyz * dz;
                const float qdz = qxz * dx + qyz * dy + qzz * dz;
                const float qd2 = dx * qdx + dy * qdy + dz * qdz;
                const float ir = __builtin_amdgcn_rsqf(dist2);
                const float ir2 = ir * ir;
                const float ir5 = ir2 * ir2 * ir;
                const float ir7 = ir5 * ir2;
                const float c1 = -ir5;
                const float c2 = 2.5f * qd2 * ir7;
                ex += c1 * qdx + c2 * dx;
                ey += c1 * qdy + c2 * dy;
                ez += c1 * qdz + c2 * dz;
            }
            kadd<KAHAN>(a.x[t], a.cx[t], ex);
            kadd<KAHAN>(a.y[t], a.cy[t], ey);
            kadd<KAHAN>(a.z[t], a.cz[t], ez);
        }
    }
}

struct WalkArgs {
    const int32_t* leaf_list;
    const int32_t* nd_start;
    const int32_t* nd_count;
    const int32_t* nd_child;
    const float4* nd_rec;  // 4 per node: com, quad, (Qyz, Qzz, first particle), aux
    const double4* nd_com64;
    const double* nd_rmax64;
    const double* nd_geom;
    const int32_t* sys_root;
    const int32_t* pos_sys;
    const float4* spos;
    const uint32_t* perm;
    const DevScalars* sc;
    float* ax;
    float* ay;
    float* az;
    int32_t* cnt_far;
    int32_t* cnt_near;
    float eps2;
    float theta;
    float theta2;
    double theta2_64;
};

template <bool QUAD, bool KAHAN, bool COUNT>
__global__ __launch_bounds__(64 * kWalkWaves) void k_walk(WalkArgs A) {
    __shared__ int s_stack[kWalkWaves][kStackCap];
    __shared__ int s_far[kWalkWaves][kFarRing];
    __shared__ int s_near[kWalkWaves][kNearRing];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = lane_id();
    int* stack = s_stack[wave];
    int* far = s_far[wave];
    int* near = s_near[wave];
    const int nl = A.sc->n_leaves;
    const float eps2 = A.eps2, theta = A.theta, theta2 = A.theta2;

    for (int li = blockIdx.x * kWalkWaves + wave; li < nl; li += gridDim.x * kWalkWaves) {
        const int g = A.leaf_list[li];
        const int gst = A.nd_start[g];
        const int gcnt = A.nd_count[g];
        const float4 gc = A.nd_rec[4 * g];
        const float gR = A.nd_rec[4 * g + 3].z;
        const double4 g64 = A.nd_com64[g];
        const double gR64 = A.nd_rmax64[g];
        const float gscale = fabsf(gc.x) + fabsf(gc.y) + fabsf(gc.z) + gR;
        const int root = A.sys_root[A.pos_sys[gst]];

        for (int t0 = 0; t0 < gcnt; t0 += kMaxT) {
            const int nt = min(kMaxT, gcnt - t0);
            float tx[kMaxT], ty[kMaxT], tz[kMaxT];
            Acc acc;
#pragma unroll
            for (int t = 0; t < kMaxT; ++t) {
                const float4 p = (t < nt) ? A.spos[gst + t0 + t] : make_float4(0.f, 0.f, 0.f, 0.f);
                tx[t] = p.x;
                ty[t] = p.y;
                tz[t] = p.z;
                acc.x[t] = acc.y[t] = acc.z[t] = 0.f;
                acc.cx[t] = acc.cy[t] = acc.cz[t] = 0.f;
            }
            int sp = 1;
            if (lane == 0) stack[0] = root;
            wave_sync();
            int far_head = 0, far_tail = 0, near_head = 0, near_tail = 0;
            int n_far = 0, n_near = 0;
            while (sp > 0) {
                const int k = (sp <= kBatchThresh) ? min(64, sp) : 1;
                int node = -1;
                if (lane < k) node = stack[sp - k + lane];
                sp -= k;
                wave_sync();
                bool is_far = false, is_near = false, is_open = false;
                int lstart = 0, lcount = 0;
                if (node >= 0 && node != g) {
                    const float4 cm = A.nd_rec[4 * node];
                    if (cm.w > 0.f) {
                        const float4 ax = A.nd_rec[4 * node + 3];
                        const uint32_t info = __float_as_uint(ax.w);
                        if (info & kLeafBit) {
                            is_near = true;
                            lstart = A.nd_start[node];
                            lcount = static_cast<int>(info & ~kLeafBit);
                        } else {
                            const float dx = cm.x - gc.x, dy = cm.y - gc.y, dz = cm.z - gc.z;
                            const float dist2 = dx * dx + dy * dy + dz * dz;
                            const float worst = sqrtf(dist2) - gR;
                            bool accept = worst > 0.f && ax.y < theta2 * worst * worst;
                            // Within FP32 reach of the boundary the decision is redone in
                            // FP64 exactly as CollectGroupInteractions writes it
                            // (Solvers.cpp:261-269), so the interaction lists are the
                            // reference's.  The slack bounds the FP32 error of worst and size.
                            const float sz = sqrtf(ax.y);
                            const float slack = 1e-4f * sz + 1e-6f * (gscale + fabsf(cm.x) + fabsf(cm.y) + fabsf(cm.z));
                            if (fabsf(theta * worst - sz) <= slack) {
                                const double4 c64 = A.nd_com64[node];
                                const double ex = __dsub_rn(c64.x, g64.x), ey = __dsub_rn(c64.y, g64.y),
                                             ez = __dsub_rn(c64.z, g64.z);
                                const double d2 =
                                    __dadd_rn(__dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey)), __dmul_rn(ez, ez));
                                const double size = 2.0 * A.nd_geom[4 * node + 3];
                                const double w64 = __dsub_rn(__dsqrt_rn(d2), gR64);
                                accept = (w64 > 0.0) &&
                                         (__dmul_rn(size, size) < __dmul_rn(__dmul_rn(A.theta2_64, w64), w64));
                            }
                            if (accept)
                                is_far = true;
                            else
                                is_open = true;
                        }
                    }
                }
                // far cells -> ring
                const uint64_t fm = ballot(is_far);
                if (is_far) far[(far_tail + mbcnt(fm)) & (kFarRing - 1)] = node;
                far_tail += __popcll(fm);
                // opened cells -> push children
                int kids[8];
                int nk = 0;
                if (is_open) {
                    const int4 c0 = reinterpret_cast<const int4*>(A.nd_child)[2 * node];
                    const int4 c1 = reinterpret_cast<const int4*>(A.nd_child)[2 * node + 1];
                    kids[0] = c0.x; kids[1] = c0.y; kids[2] = c0.z; kids[3] = c0.w;
                    kids[4] = c1.x; kids[5] = c1.y; kids[6] = c1.z; kids[7] = c1.w;
#pragma unroll
                    for (int c = 0; c < 8; ++c) nk += kids[c] >= 0;
                }
                const int kincl = wave_incl_scan(nk);
                const int ktot = rdlane(kincl, 63);
                if (is_open) {
                    int w = sp + kincl - nk;
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (kids[c] >= 0) stack[w++] = kids[c];
                }
                sp += ktot;
                // near leaves of <= ncrit particles -> ring; larger (depth-capped) ones inline
                const bool small = is_near && lcount <= kNcrit;
                const int ncnt = small ? lcount : 0;
                const int nincl = wave_incl_scan(ncnt);
                const int ntot = rdlane(nincl, 63);
                if (small) {
                    const int w = near_tail + nincl - ncnt;
                    for (int q = 0; q < lcount; ++q) near[(w + q) & (kNearRing - 1)] = lstart + q;
                }
                near_tail += ntot;
                wave_sync();
                uint64_t bm = ballot(is_near && !small);
                while (bm) {
                    const int src = __ffsll(static_cast<long long>(bm)) - 1;
                    bm &= bm - 1;
                    const int bs = rdlane(lstart, src);
                    const int bc = rdlane(lcount, src);
                    for (int q0 = 0; q0 < bc; q0 += 64) {
                        const int q = q0 + lane;
                        if (q < bc) p2p<KAHAN>(acc, tx, ty, tz, nt, A.spos[bs + q], eps2, -1);
                    }
                    if (COUNT) n_near += bc;
                }
                if (far_tail - far_head >= 64) {
                    const int nd = far[(far_head + lane) & (kFarRing - 1)];
                    m2p<QUAD, KAHAN>(acc, tx, ty, tz, nt, A.nd_rec[4 * nd], A.nd_rec[4 * nd + 1],
                                     A.nd_rec[4 * nd + 3].x, eps2);
                    far_head += 64;
                    if (COUNT) n_far += 64;
                }
                while (near_tail - near_head >= 64) {
                    const int q = near[(near_head + lane) & (kNearRing - 1)];
                    p2p<KAHAN>(acc, tx, ty, tz, nt, A.spos[q], eps2, -1);
                    near_head += 64;
                    if (COUNT) n_near += 64;
                }
                wave_sync();
            }
            // flush the rings
            {
                const int rem = far_tail - far_head;
                if (lane < rem) {
                    const int nd = far[(far_head + lane) & (kFarRing - 1)];
                    m2p<QUAD, KAHAN>(acc, tx, ty, tz, nt, A.nd_rec[4 * nd], A.nd_rec[4 * nd + 1],
                                     A.nd_rec[4 * nd + 3].x, eps2);
                }
                if (COUNT) n_far += rem;
            }
            {
                const int rem = near_tail - near_head;
                if (lane < rem) {
                    const int q = near[(near_head + lane) & (kNearRing - 1)];
                    p2p<KAHAN>(acc, tx, ty, tz, nt, A.spos[q], eps2, -1);
                }
                if (COUNT) n_near += rem;
            }
            // the group's own members, self excluded
            for (int q0 = 0; q0 < gcnt; q0 += 64) {
                const int q = q0 + lane;
                if (q < gcnt) p2p<KAHAN>(acc, tx, ty, tz, nt, A.spos[gst + q], eps2, q - t0);
            }
            if (COUNT && t0 == 0 && lane == 0) {
                A.cnt_far[g] = n_far;
                A.cnt_near[g] = n_near + gcnt - 1;
            }
            // reduce each target over the wave and write by original index
#pragma unroll
            for (int t = 0; t < kMaxT; ++t) {
                if (t < nt) {
                    float sx = acc.x[t], sy = acc.y[t], sz = acc.z[t];
                    if (KAHAN) {
                        sx -= acc.cx[t];
                        sy -= acc.cy[t];
                        sz -= acc.cz[t];
                    }
                    sx = wave_sum(sx);
                    sy = wave_sum(sy);
                    sz = wave_sum(sz);
                    if (lane == t) {
                        const uint32_t o = A.perm[gst + t0 + t];
                        A.ax[o] = sx;
                        A.ay[o] = sy;
                        A.az[o] = sz;
                    }
                }
            }
            wave_sync();
        }
    }
}

// ============================================================================ bin walk
// One wavefront per bin: <= 64 consecutive sorted particles of one system (one target
// per lane).  The bin's particles belong to <= 64 leaves; leaf j of the bin is bit j of a
// 64-bit mask.  The tree is walked once for the whole bin: every stack entry carries the
// mask of bin leaves still descending through it, and each visited node is classified for
// every active leaf with that leaf's own group MAC, exactly as CollectGroupInteractions
// classifies it for that leaf alone.  So each leaf's far and near lists are the
// reference's; a far cell or near leaf reached by several leaves of the bin is staged once
// and evaluated by all 64 lanes, each lane keeping the contribution only if its own leaf's
// bit is in the entry's mask.  Sources are read with wave-uniform (scalar) loads.
// The walk over the tree visits up to 64 stacked nodes per step, one per lane, each lane
// looping over the bin's leaves (leaf data broadcast from LDS).
constexpr int kBinWaves = 4;

constexpr int kBinStack = 640;
constexpr int kBinRing = 128;  // power of two; far and near rings

struct BinArgs {
    int n_pad;  // index of the first zero pad entry after the sorted particles
    const int32_t* bin_start;
    const int32_t* bin_count;
    int n_bins;
    const int32_t* leaf_of;
    const int32_t* nd_start;
    const int32_t* nd_child;
    const float4* nd_rec;  // 4 per node: com, quad, (Qyz, Qzz, first particle), aux
    const double4* nd_com64;
    const double* nd_rmax64;
    const double* nd_geom;
    const int32_t* sys_root;
    const int32_t* pos_sys;
    const float4* spos;
    const float4* srel;  // FP64 state: sorted positions relative to their leaf's centre
    const uint32_t* perm;
    DevScalars* sc;
    float* ax;
    float* ay;
    float* az;
    float4* acc_sorted;  // learned walk: accelerations by sorted position (the correction adds in)
    int32_t* cnt_far;
    int32_t* cnt_near;
    Soft soft;
    float inv_theta;
    double theta2_64;
    // learned correction: accepted pairs with size / (dist - R) > band and a correction
    // scale G M s^3 / r^5 >= amin are listed
    float inv_band;
    float amin;
    float inv_amin;  // 1 / amin (0 without amin)
    int2* lpairs;
    uint32_t lcap;
};

// Exclusive prefix over the wave of small non-negative counts (< 16) via bit-sliced ballots.
__device__ __forceinline__ int wave_excl_small(int v, int* total) {
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

// REL (an FP64 state): every difference a force takes is formed in FP32 from positions
// relative to FP64 points near the pair -- targets relative to the centre of their bin's
// first leaf (cb), sources relative to their own leaf's centre, cells' centres of mass
// relative to cb -- plus the FP32 rounding of the offset between those FP64 points, so no
// absolute position is rounded to FP32.  Near leaves are then evaluated leaf by leaf.
template <bool QUAD, bool KAHAN, bool COUNT, bool LEARN, int NT, bool SPL, bool REL>
__global__ __launch_bounds__(64 * kBinWaves) void k_walk_bins(BinArgs A) {
    __shared__ int s_snode[kBinWaves][kBinStack];
    __shared__ uint64_t s_smask[kBinWaves][kBinStack];
    __shared__ __attribute__((aligned(16))) int s_fnode[kBinWaves][kBinRing];
    // the groups' masks of the staged far cells and near leaves, as their low and high
    // 32-bit words in two arrays: a lane reads only the word holding its own group's bit
    __shared__ __attribute__((aligned(16))) uint32_t s_fmlo[kBinWaves][kBinRing];
    __shared__ __attribute__((aligned(16))) uint32_t s_fmhi[kBinWaves][kBinRing];
    __shared__ int2 s_qsc[kBinWaves][kBinRing];  // near leaf: first sorted particle, count
    __shared__ uint32_t s_qmlo[kBinWaves][kBinRing];
    __shared__ uint32_t s_qmhi[kBinWaves][kBinRing];
    __shared__ int s_qnd[kBinWaves][REL ? kBinRing : 1];  // REL: node of each staged near leaf
    // the bin's leaf spheres at set-up; during the traversal, the records of the nodes whose
    // groups are tested one node at a time (centre of mass, size / theta)
    __shared__ float4 s_lcom[kBinWaves][64];
    __shared__ int s_lid[kBinWaves][64];
    __shared__ float4 s_nstage[kBinWaves][NT];             // the near field's current group of sources
    __shared__ int s_next[kBinWaves];

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = lane_id();
    int* snode = s_snode[wave];
    uint64_t* smask = s_smask[wave];
    int* fnode = s_fnode[wave];
    uint32_t* fmlo = s_fmlo[wave];
    uint32_t* fmhi = s_fmhi[wave];
    int2* qsc = s_qsc[wave];
    uint32_t* qmlo = s_qmlo[wave];
    uint32_t* qmhi = s_qmhi[wave];
    int* qnd = s_qnd[wave];
    float4* lcom = s_lcom[wave];
    int* lid = s_lid[wave];
    float4* nstage = s_nstage[wave];
    const Soft soft = A.soft;
    const float inv_theta = A.inv_theta;
    // XCD-aware work distribution: workgroups are dispatched round-robin over the 8 XCDs,
    // so XCD x takes the contiguous bin range [x*per, (x+1)*per) and its waves pull bins
    // from it in order through one counter per XCD (neighbouring bins share tree nodes in
    // that XCD's L2).
    const int xcd = blockIdx.x & 7;
    const int per = (A.n_bins + 7) >> 3;
    // once its own range is drained a wave helps the other XCDs' ranges (tail balance)
    int victim = 0;
    // stack headroom: a depth-first descent pushes at most 7 siblings per level
    const int reserve = 8 * (A.sc->max_depth + 1);
    // learned selection with a band rule (learn_band > 0; without one every accepted pair is
    // in the band and only the correction scale selects)
    const bool band_on = LEARN && A.inv_band < 1e29f;

    for (;;) {
        int b = -1;
        while (victim < 8) {
            const int r = (xcd + victim) & 7;
            if (lane == 0) s_next[wave] = static_cast<int>(atomicAdd(&A.sc->work[r], 1u));
            wave_sync();
            const int k = __builtin_amdgcn_readfirstlane(s_next[wave]);
            const int lo = r * per;
            if (lo + k < min(A.n_bins, lo + per)) {
                b = lo + k;
                break;
            }
            ++victim;
        }
        if (b < 0) break;
        const int start = cld(A.bin_start, b);
        const int cnt = cld(A.bin_count, b);
        const bool tv = lane < cnt;
        const int tpos = start + (tv ? lane : 0);
        float4 tp = A.spos[tpos];
        const int mylf = tv ? A.leaf_of[tpos] : -1;
        const int prevlf = (tv && lane > 0) ? A.leaf_of[tpos - 1] : -2;
        const uint64_t first = ballot(tv && mylf != prevlf);
        const int nl = __popcll(first);
        const int j = tv ? (__popcll(first & ((2ull << lane) - 1ull)) - 1) : 63;
        // the mask word and bit of this lane's group (lanes past the bin's targets read
        // group 63's, and their sums are dropped)
        const bool jhi = j >= 32;
        const int jsh = j & 31;
        const uint64_t all = (nl >= 64) ? ~0ull : ((1ull << nl) - 1ull);
        if (tv && mylf != prevlf) {
            const float4 c = A.nd_rec[4 * mylf];
            const float4 la = A.nd_rec[4 * mylf + 3];
            lcom[j] = make_float4(c.x, c.y, c.z, la.z);
            lid[j] = mylf;
            // the leaf's own members (self excluded when evaluated)
            qsc[j] = make_int2(__float_as_int(A.nd_rec[4 * mylf + 2].z), static_cast<int>(__float_as_uint(la.w) & ~kLeafBit));
            qmlo[j] = (j < 32) ? (1u << jsh) : 0u;
            qmhi[j] = (j < 32) ? 0u : (1u << jsh);
            if (REL) qnd[j] = mylf;
        }
        const int root = cld(A.sys_root, cld(A.pos_sys, start));
        if (lane == 0) {
            snode[0] = root;
            smask[0] = all;
        }
        wave_sync();
        // lane j holds leaf group j's sphere (centre of mass, radius) and node id for the
        // leaf-major group tests (lanes past the bin's leaves hold leaf 0 and are never asked)
        const float4 my_lc = lcom[lane < nl ? lane : 0];
        const int my_lid = lid[lane < nl ? lane : 0];
        // the group's share of the FP32 test's rounding bound below: 4e-6 (2 |c_group|_1 + R)
        const float my_ub = 4e-6f * (2.f * (fabsf(my_lc.x) + fabsf(my_lc.y) + fabsf(my_lc.z)) + my_lc.w);
        // REL: the bin's reference point cb (its first leaf's FP64 centre) and each target
        // relative to it, from its offset in its own leaf plus the FP64 offset of the leaves
        double rbx = 0.0, rby = 0.0, rbz = 0.0;
        if (REL) {
            const double4 g0 = cld(reinterpret_cast<const double4*>(A.nd_geom), __builtin_amdgcn_readfirstlane(lid[0]));
            rbx = g0.x;
            rby = g0.y;
            rbz = g0.z;
            if (tv) {
                const float4 r = A.srel[tpos];
                const double4 gl = reinterpret_cast<const double4*>(A.nd_geom)[mylf];
                tp.x = static_cast<float>((gl.x - rbx) + static_cast<double>(r.x));
                tp.y = static_cast<float>((gl.y - rby) + static_cast<double>(r.y));
                tp.z = static_cast<float>((gl.z - rbz) + static_cast<double>(r.z));
            }
        }
        // leaves whose first particle lies in this bin: a leaf that starts in the previous
        // bin gets its pairs listed there (its MAC decisions are the same in both bins)
        uint64_t own_m = all;
        if (LEARN && nl > 0 && A.nd_start[lid[0]] < start) own_m &= ~1ull;
        // a sphere (cb, Rb) holding every group sphere (com_g, R_g) of the bin: a node whose
        // distance to cb clears (fails) every group's MAC by more than the FP32 rounding
        // margin is accepted (opened) for all groups at once; by the triangle inequality
        // each group's own test would decide the same.  A NaN centre disables it.
        float cbx, cby, cbz, Rb, cbscale;
        {
            const bool isfirst = tv && mylf != prevlf;
            const float4 lc = isfirst ? lcom[j] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float inv_nl = 1.f / static_cast<float>(max(nl, 1));
            cbx = wave_sum(lc.x) * inv_nl;
            cby = wave_sum(lc.y) * inv_nl;
            cbz = wave_sum(lc.z) * inv_nl;
            const float ex = lc.x - cbx, ey = lc.y - cby, ez = lc.z - cbz;
            float r = isfirst ? sqrtf(ex * ex + ey * ey + ez * ez) + lc.w : 0.f;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) r = fmaxf(r, __shfl_xor(r, o, 64));
            Rb = r * (1.f + 1e-5f);
            cbscale = fabsf(cbx) + fabsf(cby) + fabsf(cbz) + Rb;
        }

        float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
        int n_far = 0, n_near = 0;
        int sp = 1;
        int f_head = 0, f_tail = 0, q_head = 0, q_tail = nl;

        // evaluates `n` staged far cells starting at ring position `h`, four at a time
        // (their scalar loads issued before any arithmetic); every lane computes every
        // cell and keeps it only if its leaf's bit is in the cell's mask
        // (the ring is consumed from multiples of 64, so a group of four slots is one aligned
        // 16-byte node read and one 16-byte read of the lane's mask words; slots past n repeat
        // the first node with zero masks, written here into the free slots after the last)
        const uint32_t* fmw = jhi ? fmhi : fmlo;
        auto eval_far = [&](int h, int n) __attribute__((always_inline)) {
            float sx = 0.f, sy = 0.f, sz = 0.f;
            const int hb = h & (kBinRing - 1);
            if (n & 3) {
                if (lane < 4 - (n & 3)) {
                    fmlo[hb + n + lane] = 0u;
                    fmhi[hb + n + lane] = 0u;
                }
            }
            for (int e0 = 0; e0 < n; e0 += 4) {
                int nd[4];
                float add[4];
                float4 cm[4], q[4];
                float2 q2[4];
                const int4 n4 = *reinterpret_cast<const int4*>(fnode + hb + e0);
                const uint4 w4 = *reinterpret_cast<const uint4*>(fmw + hb + e0);
                const int nv[4] = {n4.x, n4.y, n4.z, n4.w};
                const uint32_t wv[4] = {w4.x, w4.y, w4.z, w4.w};
                const int nd0 = __builtin_amdgcn_readfirstlane(nv[0]);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    nd[u] = (u == 0 || e0 + u < n) ? __builtin_amdgcn_readfirstlane(nv[u]) : nd0;
                    add[u] = mask_pick(wv[u], jsh, 0.f, __builtin_inff());
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cm[u] = cld(A.nd_rec, 4 * nd[u]);
                    if (REL) {
                        // the centre of mass relative to cb: an FP64 difference, rounded once
                        const double4 c64 = cld(A.nd_com64, nd[u]);
                        cm[u].x = static_cast<float>(c64.x - rbx);
                        cm[u].y = static_cast<float>(c64.y - rby);
                        cm[u].z = static_cast<float>(c64.z - rbz);
                    }
                    q[u] = cld(A.nd_rec, 4 * nd[u] + 1);
                    {
                        const float4 w = cld(A.nd_rec, 4 * nd[u] + 2);
                        q2[u] = make_float2(w.x, w.y);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    m2p_add_r2<QUAD, SPL>(cm[u], q[u], q2[u], tp.x, tp.y, tp.z, soft, add[u], sx, sy, sz);
                    if (COUNT) n_far += add[u] == 0.f;
                }
            }
            kadd<KAHAN>(ax, cx, sx);
            kadd<KAHAN>(ay, cy, sy);
            kadd<KAHAN>(az, cz, sz);
        };
        // evaluates the particles of `n` staged near leaves starting at ring position `h`,
        // packed NT at a time across leaf boundaries (a scalar cursor walks the leaves -- NT is
        // the P2P tile, chosen by the autotuner).  Lanes 0..NT-1 load the group's NT records
        // with one per-lane load and stage them in LDS, from where every lane reads each
        // record (a wave-uniform record loaded by all 64 lanes costs the texture path 64
        // deliveries).  Unused slots of the last group read the zero pad after the sorted
        // particles.
        // FAST: every pair term is finite (Plummer with eps > 0, or the spline), so a lane
        // drops an unused source arithmetically instead of selecting, and its own particle
        // adds an exact zero (d = 0): Plummer adds +inf instead of eps^2 to its r^2 (the
        // reciprocal root of +inf is +0, so the term is +0, as a zero weight gave, one
        // multiply less per source), the spline weighs the mass by its use (1 or 0);
        // otherwise unused sources and the target itself are dropped by a select.
        auto eval_near_t = [&](int h, int n, auto fast) __attribute__((always_inline)) {
            constexpr bool FAST = decltype(fast)::value;
            float sx = 0.f, sy = 0.f, sz = 0.f;
            int e = 0, cur = 0, end = 0;  // the cursor: leaf e, its next particle cur, its end
            // a source's per-lane use: its r^2 addend (Plummer FAST) or its mass weight
            constexpr bool ADDEND = FAST && !SPL;
            const float wyes = ADDEND ? soft.eps2 : 1.f, wno = ADDEND ? __builtin_inff() : 0.f;
            float wuse = wno;
            const __amdgpu_buffer_rsrc_t srec = rec_rsrc(A.spos, A.n_pad + 1);
            // the n <= 64 leaves, one per lane (two LDS reads for all of them): the cursor takes
            // each leaf's words with v_readlane instead of waiting on an LDS read per leaf.  Past
            // the last leaf the cursor reads a sentinel (no use, an end never reached), so a
            // slot costs one test of the current leaf's end.
            const int lslot = (h + lane) & (kBinRing - 1);
            const int2 my_q = lane < n ? qsc[lslot] : make_int2(0, 0);
            const uint32_t my_qlo = lane < n ? qmlo[lslot] : 0u, my_qhi = lane < n ? qmhi[lslot] : 0u;
            auto load_entry = [&](int ee) __attribute__((always_inline)) {
                // (lane ee & 63 is read unconditionally and dropped past the last leaf)
                const int st_r = rdlane(my_q.x, ee & 63), c_r = rdlane(my_q.y, ee & 63);
                const uint32_t lo_r = rdlane(my_qlo, ee & 63), hi_r = rdlane(my_qhi, ee & 63);
                const bool live = ee < n;
                cur = live ? st_r : A.n_pad;
                end = live ? st_r + c_r : -1;
                // the word of the leaf's mask (wave-uniform) holding the lane's group bit, by
                // one 64-bit shift of 0 or 32; that bit picks the weight
                const uint64_t em = live ? ((static_cast<uint64_t>(hi_r) << 32) | lo_r) : 0ull;
                wuse = mask_pick(static_cast<uint32_t>(em >> (j & 32)), jsh, wyes, wno);
            };
            load_entry(0);
            int groups = 0;
            while (e < n) {
                int idx[NT];
                float us[NT];
                int vidx = A.n_pad;  // lane u < NT: the group's u-th source
#pragma unroll
                for (int u = 0; u < NT; ++u) {
                    idx[u] = min(cur, A.n_pad);
                    us[u] = wuse;
                    vidx = wrlane(vidx, idx[u], u);
                    if (COUNT) n_near += (wuse == wyes) && (cur != tpos);
                    if (__builtin_expect(++cur == end, 0)) load_entry(++e);
                }
                if (lane < NT) nstage[lane] = bld_lane(srec, vidx);
                wave_sync();
#pragma unroll
                for (int u = 0; u < NT; ++u) {
                    const float4 q = nstage[u];
                    const float dx = q.x - tp.x, dy = q.y - tp.y, dz = q.z - tp.z;
                    float f;
                    if (ADDEND) {
                        const float inv = __builtin_amdgcn_rsqf(fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, us[u]))));
                        f = q.w * (inv * inv * inv);
                    } else if (FAST) {
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
                if (++groups == 64 / NT) {
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
        // REL: the near leaves one at a time -- the target's offset from the leaf's centre
        // (tp minus the FP32 offset of that centre from cb) against the sources' offsets
        // from the same centre; the target itself and the pad slots are dropped by select
        auto eval_near_rel = [&](int h, int n) __attribute__((always_inline)) {
            float sx = 0.f, sy = 0.f, sz = 0.f;
            const __amdgpu_buffer_rsrc_t srec = rec_rsrc(A.srel, A.n_pad + 1);
            for (int e = 0; e < n; ++e) {
                const int slot = (h + e) & (kBinRing - 1);
                const int2 e2 = qsc[slot];
                const int st = __builtin_amdgcn_readfirstlane(e2.x);
                const int c = __builtin_amdgcn_readfirstlane(e2.y);
                const int nd = __builtin_amdgcn_readfirstlane(qnd[slot]);
                const bool use = (((jhi ? qmhi[slot] : qmlo[slot]) >> jsh) & 1u) != 0;
                const double4 g = cld(reinterpret_cast<const double4*>(A.nd_geom), nd);
                const float txr = tp.x - static_cast<float>(g.x - rbx);
                const float tyr = tp.y - static_cast<float>(g.y - rby);
                const float tzr = tp.z - static_cast<float>(g.z - rbz);
                for (int k0 = 0; k0 < c; k0 += 8) {
                    int idx[8];
                    float4 spt[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        idx[u] = (k0 + u < c) ? st + k0 + u : A.n_pad;
                        spt[u] = bld(srec, idx[u]);
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const float dx = spt[u].x - txr, dy = spt[u].y - tyr, dz = spt[u].z - tzr;
                        const bool take = use && k0 + u < c && idx[u] != tpos;
                        float f = pair_coef<SPL>(dx * dx + dy * dy + dz * dz, spt[u].w, soft);
                        f = take ? f : 0.f;
                        if (COUNT) n_near += take;
                        sx += f * dx;
                        sy += f * dy;
                        sz += f * dz;
                    }
                }
                if ((e & 7) == 7) {
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
        // learned pair listing: up to 128 pending (cell, leaf mask) entries, two per lane, are
        // written leaf-major (one ballot per leaf and row), so that the consumer finds each
        // leaf's pairs of up to 128 entries in one run and sums a run before one atomic update
        int pnode0 = 0, pnode1 = 0;
        uint64_t pm0 = 0, pm1 = 0;
        int npend = 0;
        auto flush_pairs = [&]() __attribute__((always_inline)) {
            const int tot = wave_add(__popcll(pm0) + __popcll(pm1));
            const uint64_t used = wave_or64(pm0 | pm1);
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&A.sc->learn_pairs, static_cast<uint32_t>(tot));
            uint32_t w = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(base)));
            for (uint64_t t = used; t; t &= t - 1) {
                const int jj = __ffsll(static_cast<long long>(t)) - 1;
                const int lf = rdlane(my_lid, jj);
                const bool a0 = (pm0 >> jj) & 1ull, a1 = (pm1 >> jj) & 1ull;
                const uint64_t b0 = ballot(a0), b1 = ballot(a1);
                if (a0) {
                    const uint32_t idx = w + static_cast<uint32_t>(mbcnt(b0));
                    if (idx < A.lcap) A.lpairs[idx] = make_int2(lf, pnode0);
                }
                w += static_cast<uint32_t>(__popcll(b0));
                if (a1) {
                    const uint32_t idx = w + static_cast<uint32_t>(mbcnt(b1));
                    if (idx < A.lcap) A.lpairs[idx] = make_int2(lf, pnode1);
                }
                w += static_cast<uint32_t>(__popcll(b1));
            }
            pm0 = 0;
            pm1 = 0;
            npend = 0;
        };
        const bool near_fast = SPL || soft.eps2 > 0.f;
        auto eval_near = [&](int h, int n) __attribute__((always_inline)) {
            if (REL)
                eval_near_rel(h, n);
            else if (near_fast)
                eval_near_t(h, n, std::true_type{});
            else
                eval_near_t(h, n, std::false_type{});
        };

        // pops the next batch of up to 64 stacked nodes (one per lane) and issues the loads of
        // their record words, which land while the step before evaluates its staged sources
        int node = -1;
        uint64_t act = 0;
        float4 rcm = make_float4(0.f, 0.f, 0.f, 0.f), rau = rcm;
        int rst = 0;
        auto pop = [&]() __attribute__((always_inline)) {
            const int room = kBinStack - sp - reserve;
            const int k = min(min(64, sp), max(1, room / 7));
            node = -1;
            act = 0;
            if (lane < k) {
                node = snode[sp - k + lane];
                act = smask[sp - k + lane];
            }
            sp -= k;
            wave_sync();
            if (node >= 0) {
                // the node's record line: com, aux word and (a leaf's) first particle
                rcm = A.nd_rec[4 * node];
                rau = A.nd_rec[4 * node + 3];
                rst = __float_as_int(A.nd_rec[4 * node + 2].z);
            }
        };
        bool more = sp > 0;
        if (more) pop();
        while (more) {
            uint64_t far_m = 0, near_m = 0, open_m = 0, band_m = 0;
            uint64_t todo = 0;  // groups the sphere shortcuts left undecided for this lane's node
            float n_sth = 0.f, n_sbd = 0.f, n_r2max = 0.f;
            // learned selection of a sphere-accepted node left to the groups: band membership
            // (bneed, the groups to test) and the correction-scale test (aneed)
            uint64_t bneed = 0;
            bool aneed = false;
            int lstart = 0, lcount = 0;
            float4 cm = make_float4(0.f, 0.f, 0.f, 0.f), au = cm;
            uint32_t info = 0;
            if (node >= 0) {
                cm = rcm;
                au = rau;
                const int st_node = rst;
                info = __float_as_uint(au.w);
                if (info & kLeafBit) {
                    lstart = st_node;
                    lcount = static_cast<int>(info & ~kLeafBit);
                }
            }
            {
                // a bin leaf never takes itself as a near leaf: a leaf is one of the bin's
                // iff its particles meet the bin's, and its bit is that of the target lane
                // holding its first particle inside the bin
                const bool own = (info & kLeafBit) && lstart < start + cnt && lstart + lcount > start;
                const int src = own ? max(lstart, start) - start : 0;
                const int jsrc = __shfl(j, src, 64);
                if (own) act &= ~(1ull << jsrc);
            }
            if (node >= 0) {
                if (cm.w > 0.f && act) {
                    if (info & kLeafBit) {
                        near_m = act;
                    } else {
                        // the sphere shortcuts and the FP32 leaf test below have a slack of
                        // 1e-4 relative, which covers the hardware square root (1 ulp)
                        const float size = __builtin_amdgcn_sqrtf(au.y);
                        const float sth = size * inv_theta;  // size / theta
                        // (the learned band's distance and sphere tests keep the correctly rounded
                        // root; without a band the hardware root serves, within the slacks)
                        const float sbd = (LEARN && band_on) ? sqrtf(au.y) * A.inv_band : __builtin_inff();
                        const float cscale = fabsf(cm.x) + fabsf(cm.y) + fabsf(cm.z);
                        const float bx = cm.x - cbx, by = cm.y - cby, bz = cm.z - cbz;
                        const float db2 = bx * bx + by * by + bz * bz;
                        const float db = (LEARN && band_on) ? sqrtf(db2) : __builtin_amdgcn_sqrtf(db2);
                        const float slack_b = 1e-4f * (db + Rb + sth) + 4e-6f * (cscale + cbscale);
                        uint64_t todo_all = act;
                        if (db - Rb - sth > slack_b) {
                            far_m = act;  // accepted by every group
                            todo_all = 0;
                            if (LEARN) {
                                // band membership (a heuristic selection, decided the same way
                                // when samples are collected): by the sphere when it settles
                                // every group, else by each group's distance
                                if (!band_on || sbd - db - Rb > slack_b) {
                                    band_m = act;
                                } else if (!(db - Rb - sbd > slack_b)) {
                                    bneed = act;  // tested group by group below, leaf-major
                                }
                            }
                        } else if (sth - db - Rb > slack_b) {
                            open_m = act;  // opened by every group
                            todo_all = 0;
                        }
                        // the groups left undecided are tested below, leaf-major
                        todo = todo_all;
                        n_sth = sth;
                        if (LEARN) {
                            n_sbd = sbd;
                            // the correction scale G M s^3 / r^5 reaches amin iff r^2 <= r2max =
                            // (G M s^3 / amin)^(2/5) (no bound without amin)
                            const float gms3 = cm.w * au.y * size;
                            n_r2max = A.amin > 0.f ? __builtin_amdgcn_exp2f(0.4f * __builtin_amdgcn_logf(gms3 * A.inv_amin))
                                                   : __builtin_inff();
                            if (A.amin > 0.f && (band_m | bneed)) {
                                // the scale test of a sphere-accepted node: by the bin sphere's
                                // bounds on r when they settle it, else group by group below (the
                                // undecided nodes take it inside their tests)
                                // (1e-6 relative covers the root's rounding: the bounds only
                                // widen, the groups' own tests decide what they leave)
                                const float rlo = fmaxf(db * (1.f - 1e-6f) - Rb, 0.f), rhi = db * (1.f + 1e-6f) + Rb;
                                if (rlo * rlo > n_r2max) {
                                    band_m = 0;
                                    bneed = 0;
                                } else if (rhi * rhi > n_r2max) {
                                    aneed = true;
                                }
                            }
                        }
                    }
                }
            }
            // the children of a node that some group may open, loaded before the group tests
            // decide (their latency hides behind the tests; a node every group accepts loads
            // none, and one the tests accept in the end loads them for nothing)
            int4 ch0 = make_int4(-1, -1, -1, -1), ch1 = ch0;
            if ((open_m | todo) != 0) {
                ch0 = reinterpret_cast<const int4*>(A.nd_child)[2 * node];
                ch1 = reinterpret_cast<const int4*>(A.nd_child)[2 * node + 1];
            }
            // Leaf-major group tests: one iteration per node left undecided by the sphere
            // shortcuts, each lane testing its own leaf group against that node, so that an
            // iteration serves every undecided group of the node (node-major, the wave looped
            // over the most groups any one node had left, with few lanes busy per trip).  The
            // node's record is one broadcast LDS read; the group masks are combined in scalar
            // registers, and node L's results are written into lane L with v_writelane.
            if (ballot(todo != 0 || (LEARN && (bneed != 0 || aneed)))) {
                lcom[lane] = make_float4(cm.x, cm.y, cm.z, n_sth);
                wave_sync();
            }
            for (uint64_t und = ballot(todo != 0); und; und &= und - 1) {
                const int L = __ffsll(static_cast<long long>(und)) - 1;
                const uint64_t tL = rdlane(todo, L);  // the groups to test (lanes = groups)
                const float4 nr = lcom[L];
                const float dx = nr.x - my_lc.x, dy = nr.y - my_lc.y, dz = nr.z - my_lc.z;
                const float dist2 = dx * dx + dy * dy + dz * dz;
                // worst > 0 && size^2 < theta^2 worst^2  <=>  dist > R + size/theta
                const float q = my_lc.w + nr.w;
                const float q2 = q * q;
                uint64_t acc = ballot(dist2 > q2) & tL;
                // within the FP32 test's rounding bound, 2q (1e-4 q + 2e-6 (|c_node|_1 + |c_group|_1 + R)),
                // the FP64 decision exactly as CollectGroupInteractions writes it (Solvers.cpp:261-269).
                // Only a pair with dist within 1e-4 relative of q can be decided wrongly, and for it
                // |c_node|_1 <= |c_group|_1 + sqrt(3) dist <= |c_group|_1 + 1.7338 q, so the bound
                // is taken as q (2.07e-4 q + 4e-6 (2 |c_group|_1 + R)): per-lane constants and q only
                const uint64_t closem = ballot(fabsf(dist2 - q2) <= q * fmaf(q, 2.07e-4f, my_ub)) & tL;
                if (closem) {
                    const int nodeL = rdlane(node, L);
                    bool a64 = false;
                    if ((closem >> lane) & 1ull) {
                        const double4 c64 = A.nd_com64[nodeL];
                        const double4 g64 = A.nd_com64[my_lid];
                        const double gR64 = A.nd_rmax64[my_lid];
                        const double ex = __dsub_rn(c64.x, g64.x), ey = __dsub_rn(c64.y, g64.y),
                                     ez = __dsub_rn(c64.z, g64.z);
                        const double d2 =
                            __dadd_rn(__dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey)), __dmul_rn(ez, ez));
                        const double size = 2.0 * A.nd_geom[4 * nodeL + 3];
                        const double w64 = __dsub_rn(__dsqrt_rn(d2), gR64);
                        a64 = (w64 > 0.0) && (__dmul_rn(size, size) < __dmul_rn(__dmul_rn(A.theta2_64, w64), w64));
                    }
                    acc = (acc & ~closem) | (ballot(a64) & closem);
                }
                // (an undecided node has no far or band groups before its tests)
                far_m = wrlane(far_m, acc, L);
                if (LEARN) {
                    // band and correction scale, both by distance (the node's bounds read from
                    // its lane; without a band its band distance is +inf and the scale decides)
                    const float qb = my_lc.w + rdlane(n_sbd, L);
                    band_m = wrlane(band_m, acc & ballot(dist2 <= fminf(qb * qb, rdlane(n_r2max, L))), L);
                }
            }
            open_m |= todo & ~far_m;
            if (LEARN) {
                // the learned selection of the sphere-accepted nodes left to the groups,
                // leaf-major: one iteration per such node, each lane testing its own group
                for (uint64_t und = ballot(bneed != 0 || aneed); und; und &= und - 1) {
                    const int L = __ffsll(static_cast<long long>(und)) - 1;
                    const uint64_t bnL = rdlane(bneed, L), bmL = rdlane(band_m, L);
                    const bool anL = rdlane(static_cast<int>(aneed), L) != 0;
                    const float4 nr = lcom[L];
                    const float2 lr = make_float2(rdlane(n_sbd, L), rdlane(n_r2max, L));
                    const float dx = nr.x - my_lc.x, dy = nr.y - my_lc.y, dz = nr.z - my_lc.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    const float qb = my_lc.w + lr.x;
                    uint64_t res = band_on ? (bnL & ballot(r2 <= qb * qb)) | (bmL & ~bnL) : bmL;
                    if (anL) res &= ballot(r2 <= lr.y);
                    band_m = wrlane(band_m, res, L);
                }
            }
            // far cells -> ring
            {
                const bool has = far_m != 0;
                const uint64_t bm = ballot(has);
                if (has) {
                    const int slot = (f_tail + mbcnt(bm)) & (kBinRing - 1);
                    fnode[slot] = node;
                    fmlo[slot] = static_cast<uint32_t>(far_m);
                    fmhi[slot] = static_cast<uint32_t>(far_m >> 32);
                }
                f_tail += __popcll(bm);
            }
            // learned correction: the (cell, mask of owned leaves with a band pair) entries of
            // the step join the wave's pending entries (compacted through LDS: lcom is free
            // between the group tests and the next step; written here and read back after the
            // step's one LDS sync below), which are listed leaf-major when 128 are pending and
            // at the end of the bin
            int lk = 0;
            if (LEARN) {
                const uint64_t em = band_m & own_m;
                const bool has = em != 0;
                const uint64_t bm = ballot(has);
                lk = __popcll(bm);
                if (lk) {
                    if (npend + lk > 2 * 64) flush_pairs();
                    if (has)
                        reinterpret_cast<int4*>(lcom)[mbcnt(bm)] =
                            make_int4(node, static_cast<int>(static_cast<uint32_t>(em)),
                                      static_cast<int>(static_cast<uint32_t>(em >> 32)), 0);
                }
            }
            // near leaves -> ring
            {
                const bool has = near_m != 0;
                const uint64_t bm = ballot(has);
                if (has) {
                    const int slot = (q_tail + mbcnt(bm)) & (kBinRing - 1);
                    qsc[slot] = make_int2(lstart, lcount);
                    qmlo[slot] = static_cast<uint32_t>(near_m);
                    qmhi[slot] = static_cast<uint32_t>(near_m >> 32);
                    if (REL) qnd[slot] = node;
                }
                q_tail += __popcll(bm);
            }
            // opened cells -> push children with the opening leaves' mask
            {
                const int kids[8] = {ch0.x, ch0.y, ch0.z, ch0.w, ch1.x, ch1.y, ch1.z, ch1.w};
                int nk = 0;
                if (open_m) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) nk += kids[c] >= 0;
                }
                int ktot;
                const int koff = wave_excl_small(nk, &ktot);
                if (open_m) {
                    int w = sp + koff;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        if (kids[c] >= 0) {
                            snode[w] = kids[c];
                            smask[w] = open_m;
                            ++w;
                        }
                    }
                }
                sp += ktot;
            }
            wave_sync();
            if (LEARN && lk) {
                // entry e of the pending list sits in lane e % 64 of row e / 64 (the entries in
                // lcom are read before the next step writes it: one wave's LDS accesses complete
                // in order)
                const int4* scr = reinterpret_cast<const int4*>(lcom);
                const int e0 = lane - npend, e1 = lane + 64 - npend;
                if (e0 >= 0 && e0 < lk) {
                    const int4 v = scr[e0];
                    pnode0 = v.x;
                    pm0 = (static_cast<uint64_t>(static_cast<uint32_t>(v.z)) << 32) | static_cast<uint32_t>(v.y);
                }
                if (e1 >= 0 && e1 < lk) {
                    const int4 v = scr[e1];
                    pnode1 = v.x;
                    pm1 = (static_cast<uint64_t>(static_cast<uint32_t>(v.z)) << 32) | static_cast<uint32_t>(v.y);
                }
                npend += lk;
            }
            more = sp > 0;
            if (more) pop();
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
        if (LEARN && npend > 0) flush_pairs();
        if (tv) {
            const float rx = KAHAN ? ax - cx : ax, ry = KAHAN ? ay - cy : ay, rz = KAHAN ? az - cz : az;
            if (LEARN) {
                // coalesced; the correction's apply pass scatters to input order once
                A.acc_sorted[tpos] = make_float4(rx, ry, rz, 0.f);
            } else {
                const uint32_t o = A.perm[tpos];
                A.ax[o] = rx;
                A.ay[o] = ry;
                A.az[o] = rz;
            }
            if (COUNT) {
                A.cnt_far[mylf] = n_far;
                A.cnt_near[mylf] = n_near;
            }
        }
        wave_sync();
    }
}

void launch_walk_leaf(nbg_ctx* c, const nbg_params& p) {
    WalkArgs A;
    A.leaf_list = c->leaf_list;
    A.nd_start = c->nd_start;
    A.nd_count = c->nd_count;
    A.nd_child = c->nd_child;
    A.nd_rec = c->nd_rec;
    A.sys_root = c->sys_root;
    A.pos_sys = c->pos_sys;
    A.spos = c->spos;
        A.perm = c->perm;
    A.sc = c->dsc;
    A.ax = c->ax;
    A.ay = c->ay;
    A.az = c->az;
    A.cnt_far = c->cnt_far;
    A.cnt_near = c->cnt_near;
    A.nd_com64 = c->nd_com64;
    A.nd_rmax64 = c->nd_rmax64;
    A.nd_geom = c->nd_geom;
    A.eps2 = static_cast<float>(p.eps2);
    A.theta = static_cast<float>(p.theta);
    A.theta2 = static_cast<float>(p.theta * p.theta);
    A.theta2_64 = p.theta * p.theta;
    const bool quad = p.multipole == NBG_QUADRUPOLE;
    const bool kahan = (p.flags & NBG_FLAG_KAHAN) != 0;
    const bool count = (p.flags & NBG_FLAG_COUNT) != 0 && c->cnt_far != nullptr;
    const int blocks = 256 * 8;
    const dim3 grid(blocks), block(64 * kWalkWaves);
#define NBG_WALK(Q, K, C) hipLaunchKernelGGL((k_walk<Q, K, C>), grid, block, 0, c->stream, A)
    if (quad) {
        if (kahan) {
            if (count) NBG_WALK(true, true, true); else NBG_WALK(true, true, false);
        } else {
            if (count) NBG_WALK(true, false, true); else NBG_WALK(true, false, false);
        }
    } else {
        if (kahan) {
            if (count) NBG_WALK(false, true, true); else NBG_WALK(false, true, false);
        } else {
            if (count) NBG_WALK(false, false, true); else NBG_WALK(false, false, false);
        }
    }
#undef NBG_WALK
}

template <bool SPL>
void dispatch_walk_bins(nbg_ctx* c, const BinArgs& A, bool quad, bool kahan, bool count, bool learn) {
    // 8 XCDs x 32 CUs x 3 resident blocks
    const dim3 grid(8 * 32 * 3), block(64 * kBinWaves);
#define NBG_WALKB(Q, K, C, L, T) hipLaunchKernelGGL((k_walk_bins<Q, K, C, L, T, SPL, false>), grid, block, 0, c->stream, A)
#define NBG_WALKR(Q, K, C) hipLaunchKernelGGL((k_walk_bins<Q, K, C, false, 8, SPL, true>), grid, block, 0, c->stream, A)
    const int nt = c->tune.near_tile;
    if (c->f64_state && !learn) {
        // an FP64 state: the relative-difference walk (its near field does not take the tile)
        if (quad) {
            if (count) { if (kahan) NBG_WALKR(true, true, true); else NBG_WALKR(true, false, true); }
            else { if (kahan) NBG_WALKR(true, true, false); else NBG_WALKR(true, false, false); }
        } else {
            if (count) { if (kahan) NBG_WALKR(false, true, true); else NBG_WALKR(false, false, true); }
            else { if (kahan) NBG_WALKR(false, true, false); else NBG_WALKR(false, false, false); }
        }
        return;
    }
    if constexpr (!SPL) {
        if (learn) {
            // the learned correction is defined against the quadrupole expansion
            if (kahan) {
                if (count)
                    NBG_WALKB(true, true, true, true, 8);
                else if (nt == 16)
                    NBG_WALKB(true, true, false, true, 16);
                else
                    NBG_WALKB(true, true, false, true, 8);
            } else {
                if (count) NBG_WALKB(true, false, true, true, 8); else NBG_WALKB(true, false, false, true, 8);
            }
            return;
        }
    }
    if (count) {
        if (quad) {
            if (kahan) NBG_WALKB(true, true, true, false, 8); else NBG_WALKB(true, false, true, false, 8);
        } else {
            if (kahan) NBG_WALKB(false, true, true, false, 8); else NBG_WALKB(false, false, true, false, 8);
        }
    } else if (kahan) {
        // the production variants take the tuned P2P tile
        if (quad) {
            if (nt == 4) NBG_WALKB(true, true, false, false, 4);
            else if (nt == 16) NBG_WALKB(true, true, false, false, 16);
            else NBG_WALKB(true, true, false, false, 8);
        } else {
            if (nt == 4) NBG_WALKB(false, true, false, false, 4);
            else if (nt == 16) NBG_WALKB(false, true, false, false, 16);
            else NBG_WALKB(false, true, false, false, 8);
        }
    } else {
        if (quad) NBG_WALKB(true, false, false, false, 8); else NBG_WALKB(false, false, false, false, 8);
    }
#undef NBG_WALKB
#undef NBG_WALKR
}

void launch_walk_bins(nbg_ctx* c, const nbg_params& p) {
    BinArgs A;
    A.n_pad = c->n;
    A.bin_start = c->bin_start;
    A.bin_count = c->bin_count;
    A.n_bins = c->n_bins;
    A.leaf_of = c->leaf_of;
    A.nd_start = c->nd_start;
    A.nd_child = c->nd_child;
    A.nd_rec = c->nd_rec;
    A.nd_com64 = c->nd_com64;
    A.nd_rmax64 = c->nd_rmax64;
    A.nd_geom = c->nd_geom;
    A.sys_root = c->sys_root;
    A.pos_sys = c->pos_sys;
    A.spos = c->spos;
    A.srel = c->srel;
    A.perm = c->perm;
    A.sc = c->dsc;
    A.ax = c->ax;
    A.ay = c->ay;
    A.az = c->az;
    A.acc_sorted = c->stage4;
    A.cnt_far = c->cnt_far;
    A.cnt_near = c->cnt_near;
    A.soft = make_soft(p);
    A.inv_theta = (p.theta > 0.0) ? static_cast<float>(1.0 / p.theta) : __builtin_inff();
    A.theta2_64 = p.theta * p.theta;
    const bool quad = p.multipole == NBG_QUADRUPOLE;
    const bool kahan = (p.flags & NBG_FLAG_KAHAN) != 0;
    const bool count = (p.flags & NBG_FLAG_COUNT) != 0 && c->cnt_far != nullptr;
    // (validation guarantees quadrupoles and the learned buffers; the evaluation's apply
    // pass relies on this walk writing the sorted accelerations)
    const bool learn = (p.flags & NBG_FLAG_LEARNED) != 0;
    A.inv_band = (p.learn_band > 0.0) ? static_cast<float>(1.0 / p.learn_band) : 1e30f;
    A.amin = static_cast<float>(p.learn_amin);
    A.inv_amin = p.learn_amin > 0.0 ? static_cast<float>(1.0 / p.learn_amin) : 0.f;
    A.lpairs = c->lpairs;
    A.lcap = static_cast<uint32_t>(c->lpairs_cap);
    if (p.softening == NBG_SOFT_SPLINE)
        dispatch_walk_bins<true>(c, A, quad, kahan, count, false);
    else
        dispatch_walk_bins<false>(c, A, quad, kahan, count, learn);
}

void launch_walk(nbg_ctx* c, const nbg_params& p) {
    if ((p.flags & NBG_FLAG_PER_PARTICLE) || p.mac == NBG_MAC_RELATIVE) {
        launch_walk_pp(c, p, nullptr, nullptr, nullptr);
        return;
    }
    // the leaf-group walk has no learned-pair listing and Plummer softening only
    if (c->tune.walk_kernel == 1 && !(p.flags & NBG_FLAG_LEARNED) && p.softening == NBG_SOFT_PLUMMER)
        launch_walk_leaf(c, p);
    else
        launch_walk_bins(c, p);
}

// ============================================================================ direct sum
// One target per lane; every source of the target's system streamed with wave-uniform
// (scalar) loads, so the source needs no LDS staging and each load feeds 64 lanes.
// Blocks never straddle systems (a sort block of 2048 keys = 8 direct blocks of 256).
// ComputeAccelDirect (projects/nbody_core/src/Solvers.cpp:77-106): j == i skipped.

template <bool KAHAN, int U, bool SPL>
__global__ __launch_bounds__(256) void k_direct(const float4* __restrict__ src, const int32_t* __restrict__ blk_sys,
                                                const int32_t* __restrict__ blk_start,
                                                const int32_t* __restrict__ blk_count,
                                                const int32_t* __restrict__ sys_off,
                                                const int32_t* __restrict__ pos_sys, const int32_t* __restrict__ tlist,
                                                int nt_list, const uint32_t* __restrict__ nt_dev, Soft soft,
                                                float* __restrict__ ax, float* __restrict__ ay,
                                                float* __restrict__ az, float* __restrict__ out3) {
    if (tlist) {
        // explicit target list (ComputeAccelDirectTargets, Solvers.cpp:108-134): one target
        // per lane, any systems (sources = the target's own system); the count is nt_list,
        // or *nt_dev when that is given.  Results go to out3[3 slot ..] or, without out3,
        // to the context's accelerations of the target.
        const int slot = blockIdx.x * 256 + threadIdx.x;
        const int cnt = nt_dev ? static_cast<int>(*nt_dev) : nt_list;
        if (slot >= cnt) return;
        const int i = tlist[slot];
        // a system's particles occupy the same index range in original and sorted order
        const int tsys = pos_sys[i];
        const int lo_s = sys_off[tsys], hi_s = sys_off[tsys + 1];
        const float4 t = src[i];
        float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
        for (int j = lo_s; j < hi_s; ++j) {
            if (j == i) continue;
            const float4 q = src[j];
            const float dx = q.x - t.x, dy = q.y - t.y, dz = q.z - t.z;
            const float f = pair_coef<SPL>(dx * dx + dy * dy + dz * dz, q.w, soft);
            kadd<true>(sx, cx, f * dx);
            kadd<true>(sy, cy, f * dy);
            kadd<true>(sz, cz, f * dz);
        }
        if (out3) {
            out3[3 * slot + 0] = sx - cx;
            out3[3 * slot + 1] = sy - cy;
            out3[3 * slot + 2] = sz - cz;
        } else {
            ax[i] = sx - cx;
            ay[i] = sy - cy;
            az[i] = sz - cz;
        }
        return;
    }
    // blocks of 2048 targets (the sort tiles of the context's layout); without that layout
    // (blk_sys NULL) one system of nt_list particles cut into consecutive tiles
    const int sb = blockIdx.x >> 3;
    const int off = (blockIdx.x & 7) * 256 + threadIdx.x;
    const bool layout = blk_sys != nullptr;
    const int sys = layout ? blk_sys[sb] : 0;
    const int s0 = layout ? sys_off[sys] : 0;
    const int s1 = layout ? sys_off[sys + 1] : nt_list;
    const int bstart = layout ? blk_start[sb] : sb * kSortTile;
    const bool valid = off < (layout ? blk_count[sb] : min(kSortTile, nt_list - bstart));
    const int i = bstart + off;
    const float4 t = valid ? src[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    const int sbeg = __builtin_amdgcn_readfirstlane(s0);
    const int send = __builtin_amdgcn_readfirstlane(s1);
    // U sources per step (the direct P2P tile, chosen by the autotuner): their scalar
    // loads are issued together before the arithmetic
    int j = sbeg;
    for (; j + U <= send; j += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float4 q = src[j + u];
            const float dx = q.x - t.x, dy = q.y - t.y, dz = q.z - t.z;
            float f = pair_coef<SPL>(dx * dx + dy * dy + dz * dz, q.w, soft);
            f = (j + u == i) ? 0.f : f;
            kadd<KAHAN>(sx, cx, f * dx);
            kadd<KAHAN>(sy, cy, f * dy);
            kadd<KAHAN>(sz, cz, f * dz);
        }
    }
    for (; j < send; ++j) {
        const float4 q = src[j];
        const float dx = q.x - t.x, dy = q.y - t.y, dz = q.z - t.z;
        float f = pair_coef<SPL>(dx * dx + dy * dy + dz * dz, q.w, soft);
        f = (j == i) ? 0.f : f;
        kadd<KAHAN>(sx, cx, f * dx);
        kadd<KAHAN>(sy, cy, f * dy);
        kadd<KAHAN>(sz, cz, f * dz);
    }
    if (valid) {
        if (KAHAN) {
            sx -= cx;
            sy -= cy;
            sz -= cz;
        }
        ax[i] = sx;
        ay[i] = sy;
        az[i] = sz;
    }
}

__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ x, const float* __restrict__ y,
                                              const float* __restrict__ z, const float* __restrict__ m, int n, float G,
                                              float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(x[i], y[i], z[i], G * m[i]);
}

template <bool SPL>
void dispatch_direct(nbg_ctx* c, const nbg_params& p) {
    const dim3 grid(c->n_blocks * 8), block(256);
    const int u = c->tune.direct_tile;
    const Soft soft = make_soft(p);
#define NBG_DIRECT(K, U)                                                                                          \
    hipLaunchKernelGGL((k_direct<K, U, SPL>), grid, block, 0, c->stream, c->spos, c->blk_sys, c->blk_start,        \
                       c->blk_count, c->sys_off, nullptr, nullptr, 0, nullptr, soft, c->ax, c->ay, c->az, nullptr)
    if (p.flags & NBG_FLAG_KAHAN) {
        if (u == 2) NBG_DIRECT(true, 2);
        else if (u == 8) NBG_DIRECT(true, 8);
        else NBG_DIRECT(true, 4);
    } else {
        if (u == 2) NBG_DIRECT(false, 2);
        else if (u == 8) NBG_DIRECT(false, 8);
        else NBG_DIRECT(false, 4);
    }
#undef NBG_DIRECT
}

void launch_pack(nbg_ctx* c, double G) {
    hipLaunchKernelGGL(k_pack, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->x, c->y, c->z, c->m, c->n,
                       static_cast<float>(G), c->spos);
}

void launch_direct(nbg_ctx* c, const nbg_params& p) {
    launch_pack(c, p.G);
    if (p.softening == NBG_SOFT_SPLINE)
        dispatch_direct<true>(c, p);
    else
        dispatch_direct<false>(c, p);
}

// One system of n particles given as device arrays: packed into `stage` (n float4
// records), then the direct sum with the context path's default tile (4 sources per step)
// into ax, ay, az on `stream`.
void launch_direct_raw(hipStream_t stream, const nbg_params& p, int n, const float* x, const float* y,
                       const float* z, const float* m, float4* stage, float* ax, float* ay, float* az) {
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, stream, x, y, z, m, n, static_cast<float>(p.G),
                       stage);
    const dim3 grid(((n + kSortTile - 1) / kSortTile) * 8), block(256);
    const Soft soft = make_soft(p);
    const bool kahan = (p.flags & NBG_FLAG_KAHAN) != 0;
#define NBG_DIRECT_RAW(K, S)                                                                                       \
    hipLaunchKernelGGL((k_direct<K, 4, S>), grid, block, 0, stream, stage, nullptr, nullptr, nullptr, nullptr,     \
                       nullptr, nullptr, n, nullptr, soft, ax, ay, az, nullptr)
    if (p.softening == NBG_SOFT_SPLINE) {
        if (kahan)
            NBG_DIRECT_RAW(true, true);
        else
            NBG_DIRECT_RAW(false, true);
    } else {
        if (kahan)
            NBG_DIRECT_RAW(true, false);
        else
            NBG_DIRECT_RAW(false, false);
    }
#undef NBG_DIRECT_RAW
}

void launch_direct_targets(nbg_ctx* c, const nbg_params& p, int64_t nt, const int32_t* targets,
                           const uint32_t* nt_dev, float* out3) {
    // sources in original order (the tree's sorted copy is not needed)
    launch_pack(c, p.G);
    const Soft soft = make_soft(p);
    const dim3 grid(grid_for(nt, 256)), block(256);
    if (p.softening == NBG_SOFT_SPLINE)
        hipLaunchKernelGGL((k_direct<true, 4, true>), grid, block, 0, c->stream, c->spos, c->blk_sys, c->blk_start,
                           c->blk_count, c->sys_off, c->pos_sys, targets, static_cast<int>(nt), nt_dev, soft, c->ax, c->ay, c->az,
                           out3);
    else
        hipLaunchKernelGGL((k_direct<true, 4, false>), grid, block, 0, c->stream, c->spos, c->blk_sys, c->blk_start,
                           c->blk_count, c->sys_off, c->pos_sys, targets, static_cast<int>(nt), nt_dev, soft, c->ax, c->ay, c->az,
                           out3);
}

}  // namespace nbg
