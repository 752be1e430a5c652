// nbg_tree.hip -- bounding box, reference-consistent Morton keys, LDS radix sort,
// bottom-up octree build and multipole upsweep, for gfx950.
//
// The octree reproduces AdaptiveTree<3> (projects/nbody_core/src/Tree.cpp:33-203) bit for
// bit without its serial insertion:
//  * Root box: Tree.cpp:54-88 (FP64 min/max of the positions, half = max half-extent,
//    1.0 if degenerate, padded by 1.001).
//  * Octant of a particle in a cell: p >= centre per axis (Tree.cpp:11-20); child centre
//    = parent centre +- 0.5*parent half (Tree.cpp:96-99).  Replaying that FP64 recurrence
//    per particle gives its octant digit at every level, i.e. a Morton key whose prefixes
//    are exactly the reference's cells.
//  * Topology: insertion splits a cell once it would hold ncrit+1 = 9 particles and its
//    depth is < 40 (Tree.cpp:124-135), and creates a child only for an octant that
//    receives a particle.  So a cell is a node iff its parent holds > 8 particles, and
//    it is a leaf iff it holds <= 8 (or sits at depth 40).  On sorted keys every cell is
//    a contiguous run, so each sorted position finds its leaf depth from its +-8
//    neighbours and the nodes it starts from one prefix sum, with no atomics.
//  * Numbering: nodes started at a position, in increasing depth, over positions in key
//    order, is DFS pre-order with ascending octant -- the order Finalize walks
//    (Tree.cpp:157-175).
//  * `order`: leaves in key order, members in ascending original index (insertion fills
//    buckets in index order), so each leaf's <= 8 members are re-sorted by index.
#include "nbg_common.hpp"
#include "nbg_internal.hpp"

namespace nbg {

// ============================================================================ utilities

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_i32(nbg_ctx* c, int32_t* p, int64_t n, int32_t v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, p, n, v);
}

// ------------------------------------------------------------------- generic prefix scan
// Exclusive scan in three kernels: block sums, one-block scan of the sums, block-local scan
// plus offset.  2048 elements per block (256 threads x 8, blocked per thread).
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T>
__device__ T block_excl_scan(T v, T* s_wave, T* total) {
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    T incl = wave_incl_scan(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    T wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const T s = s_wave[w];
        if (w < wave) wave_off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wave_off + incl - v;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T* __restrict__ in, int64_t n, T* __restrict__ part) {
    __shared__ T s_wave[kScanThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    T sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) sum += in[base + k];
    T tot;
    (void)block_excl_scan(sum, s_wave, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// One workgroup scans the block sums in place (exclusive) and stores the grand total.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(T* __restrict__ part, int64_t nparts, T* __restrict__ total_out) {
    __shared__ T s_wave[kScanThreads / 64];
    T carry = 0;
    for (int64_t base = 0; base < nparts; base += kScanTile) {
        const int64_t b = base + threadIdx.x * kScanItems;
        T v[kScanItems];
        T sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (b + k < nparts) ? part[b + k] : T(0);
            sum += v[k];
        }
        T tot;
        T off = block_excl_scan(sum, s_wave, &tot) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (b + k < nparts) part[b + k] = off;
            off += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const T* __restrict__ in, T* __restrict__ out, int64_t n,
                                                            const T* __restrict__ part) {
    __shared__ T s_wave[kScanThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : T(0);
        sum += v[k];
    }
    T tot;
    T off = block_excl_scan(sum, s_wave, &tot) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = off;
        off += v[k];
    }
}

// The down pass for a modest number of tiles: each block sums the tile totals before it
// itself (at most kScanInline of them), so the separate partials pass is not needed; the
// last block stores the grand total.
constexpr int kScanInline = 4096;

// `per`: partials per tile (1: k_scan_reduce's tile sums; kScanTile / 256: the sums of
// 256-element blocks a producer kernel wrote, saving the reduce launch).
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down_inline(const T* __restrict__ in, T* __restrict__ out,
                                                                   int64_t n, const T* __restrict__ part,
                                                                   T* __restrict__ total_out, int per) {
    __shared__ T s_wave[kScanThreads / 64];
    T pre = 0;
    // eight loads in flight per thread and round
    const int np = per * static_cast<int>(blockIdx.x);
    for (int i0 = 0; i0 < np; i0 += 8 * kScanThreads) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * kScanThreads + static_cast<int>(threadIdx.x);
            v[u] = (i < np) ? part[min(i, np - 1)] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) pre += v[u];
    }
    T before;
    (void)block_excl_scan(pre, s_wave, &before);
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : T(0);
        sum += v[k];
    }
    T tot;
    T off = block_excl_scan(sum, s_wave, &tot) + before;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = off;
        off += v[k];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = before + tot;
}

// part must hold grid_for(n, kScanTile) elements.  In place (in == out) is allowed.
template <typename T>
void excl_scan(hipStream_t st, const T* in, T* out, int64_t n, T* part, T* total_out) {
    const int nb = grid_for(n, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nb), dim3(kScanThreads), 0, st, in, n, part);
    if (nb <= kScanInline) {
        hipLaunchKernelGGL(k_scan_down_inline<T>, dim3(nb), dim3(kScanThreads), 0, st, in, out, n, part, total_out, 1);
        return;
    }
    hipLaunchKernelGGL(k_scan_partials<T>, dim3(1), dim3(kScanThreads), 0, st, part, static_cast<int64_t>(nb), total_out);
    hipLaunchKernelGGL(k_scan_down<T>, dim3(nb), dim3(kScanThreads), 0, st, in, out, n, part);
}

// ============================================================================ bbox + keys

// Resets the per-call device scalars (threads 0..63 of one block).
__device__ __forceinline__ void reset_scalars(DevScalars* __restrict__ sc) {
    const int t = threadIdx.x;
    if (t < 64) {
        sc->depth_count[t] = 0;
        sc->depth_offset[t] = 0;
        sc->depth_fill[t] = 0;
    }
    if (t == 0) {
        sc->n_nodes = 0;
        sc->n_leaves = 0;
        sc->max_depth = 0;
        sc->overflow = 0;
        sc->cap_exceeded = 0;
    }
    if (t < 8) sc->work[t] = 0;
    if (t == 0) {
        sc->learn_pairs = 0;
        sc->learn_fix_over = 0;
        sc->pp_m2p = 0;
        sc->pp_p2p = 0;
        sc->work2 = 0;
        sc->tcount = 0;
        sc->pp_cost_prev = sc->pp_cost_sum;
        sc->pp_cost_sum = 0;
    }
}

// Order-preserving map of FP64 values to unsigned 64-bit integers (atomic min/max).
__device__ __forceinline__ uint64_t d2ord(double d) {
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(d));
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double ord2d(uint64_t u) {
    const uint64_t b = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    return __longlong_as_double(static_cast<long long>(b));
}
// the order-preserving map of a position type and its inverse, widened to FP64
template <typename T>
struct PosOrd;
template <>
struct PosOrd<float> {
    typedef uint32_t U;
    static __device__ __forceinline__ U enc(float v) { return f2ord(v); }
    static __device__ __forceinline__ double dec(U u) { return static_cast<double>(ord2f(u)); }
    static constexpr U lo0 = 0xffffffffu, hi0 = 0u;
};
template <>
struct PosOrd<double> {
    typedef unsigned long long U;
    static __device__ __forceinline__ U enc(double v) { return d2ord(v); }
    static __device__ __forceinline__ double dec(U u) { return ord2d(u); }
    static constexpr U lo0 = ~0ull, hi0 = 0ull;
};

// Per sort block (a slice of one system): exact min/max of the positions (FP32, or FP64
// for an FP64 state), written as the block's partial (k_root reduces a system's blocks:
// hundreds of same-address atomics from every block of a launch cost more than the loads).
template <typename T>
__global__ __launch_bounds__(kSortThreads) void k_bbox(const T* __restrict__ x, const T* __restrict__ y,
                                              const T* __restrict__ z, const int32_t* __restrict__ blk_start,
                                              const int32_t* __restrict__ blk_count,
                                              typename PosOrd<T>::U* __restrict__ blk_bbox) {
    typedef typename PosOrd<T>::U U;
    __shared__ U s_red[6][4];
    const int b = blockIdx.x;
    const int start = blk_start[b], cnt = blk_count[b];
    U lo[3] = {PosOrd<T>::lo0, PosOrd<T>::lo0, PosOrd<T>::lo0}, hi[3] = {PosOrd<T>::hi0, PosOrd<T>::hi0, PosOrd<T>::hi0};
    if (cnt > 0) {
        // a block holds at most kSortItems positions per thread: all loads are issued
        // before the first is used (one memory round trip), those past the block's end
        // repeat its last position, which leaves min and max unchanged
        T px[kSortItems], py[kSortItems], pz[kSortItems];
#pragma unroll
        for (int u = 0; u < kSortItems; ++u) {
            const int i = start + min(static_cast<int>(threadIdx.x) + u * kSortThreads, cnt - 1);
            px[u] = x[i];
            py[u] = y[i];
            pz[u] = z[i];
        }
#pragma unroll
        for (int u = 0; u < kSortItems; ++u) {
            const U v[3] = {PosOrd<T>::enc(px[u]), PosOrd<T>::enc(py[u]), PosOrd<T>::enc(pz[u])};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], v[a]);
                hi[a] = max(hi[a], v[a]);
            }
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], static_cast<U>(__shfl_xor(lo[a], o, 64)));
            hi[a] = max(hi[a], static_cast<U>(__shfl_xor(hi[a], o, 64)));
        }
    }
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s_red[a][wave] = lo[a];
            s_red[3 + a][wave] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        U l = s_red[a][0], h = s_red[3 + a][0];
        for (int w = 1; w < 4; ++w) {
            l = min(l, s_red[a][w]);
            h = max(h, s_red[3 + a][w]);
        }
        blk_bbox[6 * b + a] = l;  // an empty block leaves the neutral bounds
        blk_bbox[6 * b + 3 + a] = h;
    }
}

// Root geometry per system, FP64, as Tree.cpp:54-88: one block per system reduces the
// bounds of its sort blocks.  Block 0 also resets the evaluation's device scalars (nothing
// before this kernel reads them; a launch of its own cost a few microseconds).
template <typename T>
__global__ __launch_bounds__(256) void k_root(const typename PosOrd<T>::U* __restrict__ blk_bbox,
                                              const int32_t* __restrict__ sys_blk0,
                                              const int32_t* __restrict__ sys_off, double* __restrict__ geom,
                                              DevScalars* __restrict__ sc) {
    typedef typename PosOrd<T>::U U;
    __shared__ U s_red[6][4];
    const int s = blockIdx.x;
    if (s == 0) reset_scalars(sc);
    const int b0 = sys_blk0[s], b1 = sys_blk0[s + 1];
    U lo[3] = {PosOrd<T>::lo0, PosOrd<T>::lo0, PosOrd<T>::lo0}, hi[3] = {PosOrd<T>::hi0, PosOrd<T>::hi0, PosOrd<T>::hi0};
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], blk_bbox[6 * b + a]);
            hi[a] = max(hi[a], blk_bbox[6 * b + 3 + a]);
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], static_cast<U>(__shfl_xor(lo[a], o, 64)));
            hi[a] = max(hi[a], static_cast<U>(__shfl_xor(hi[a], o, 64)));
        }
    }
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s_red[a][wave] = lo[a];
            s_red[3 + a][wave] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double dlo[3], dhi[3];
    const bool empty = sys_off[s + 1] == sys_off[s];
    for (int a = 0; a < 3; ++a) {
        U l = s_red[a][0], h = s_red[3 + a][0];
        for (int w = 1; w < 4; ++w) {
            l = min(l, s_red[a][w]);
            h = max(h, s_red[3 + a][w]);
        }
        dlo[a] = empty ? 0.0 : PosOrd<T>::dec(l);
        dhi[a] = empty ? 0.0 : PosOrd<T>::dec(h);
    }
    double half = 0.5 * (dhi[0] - dlo[0]);
    half = fmax(half, 0.5 * (dhi[1] - dlo[1]));
    half = fmax(half, 0.5 * (dhi[2] - dlo[2]));
    if (!(half > 1e-12)) half = 1.0;
    for (int a = 0; a < 3; ++a) geom[4 * s + a] = 0.5 * (dlo[a] + dhi[a]);
    geom[4 * s + 3] = half * 1.001;
}

// Morton key of every particle from the FP64 centre recurrence (see file header).
// LEVELS = 21 fills the short key; LEVELS = 40 also fills the low word; LEVELS = 16 fills
// levels 1..16 of the short key (at their places in it, levels 17..21 zero).
// pos4 (FP32 states; null otherwise): the input-order records (x, y, z, m) the per-leaf
// gather reads, written here from the positions this kernel reads anyway.
template <int LEVELS, typename T>
__global__ __launch_bounds__(256) void k_keys(const T* __restrict__ x, const T* __restrict__ y,
                                              const T* __restrict__ z, const int32_t* __restrict__ pos_sys,
                                              const double* __restrict__ geom, int n, uint64_t* __restrict__ key,
                                              uint64_t* __restrict__ klo, uint32_t* __restrict__ val,
                                              const float* __restrict__ m, float4* __restrict__ pos4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (pos4) pos4[i] = make_float4(static_cast<float>(x[i]), static_cast<float>(y[i]), static_cast<float>(z[i]), m[i]);
    const int s = pos_sys[i];
    double cx = geom[4 * s + 0], cy = geom[4 * s + 1], cz = geom[4 * s + 2], h = geom[4 * s + 3];
    const double px = static_cast<double>(x[i]), py = static_cast<double>(y[i]), pz = static_cast<double>(z[i]);
    uint64_t hi = 0, lo = 0;
#pragma unroll
    for (int l = 1; l <= LEVELS; ++l) {
        const int ox = px >= cx, oy = py >= cy, oz = pz >= cz;
        h = 0.5 * h;
        cx = cx + (ox ? h : -h);
        cy = cy + (oy ? h : -h);
        cz = cz + (oz ? h : -h);
        const uint64_t oct = static_cast<uint64_t>(ox | (oy << 1) | (oz << 2));
        if (l <= 21)
            hi = (hi << 3) | oct;
        else
            lo = (lo << 3) | oct;
    }
    if (LEVELS < 21) hi <<= 3 * (21 - LEVELS);
    key[i] = hi;
    if (LEVELS > 21) klo[i] = lo;
    val[i] = static_cast<uint32_t>(i);
}

void launch_bbox_and_keys(nbg_ctx* c, int levels) {
    hipStream_t st = c->stream;
    const bool f64 = c->f64_state;
    // the root box and the keys from the positions of the state: FP32, or FP64 (Tree.cpp's
    // own precision) for an FP64 state
    if (f64) {
        unsigned long long* bb = reinterpret_cast<unsigned long long*>(c->blk_bbox64);
        hipLaunchKernelGGL(k_bbox<double>, dim3(c->n_blocks), dim3(kSortThreads), 0, st, c->x64, c->y64, c->z64,
                           c->blk_start, c->blk_count, bb);
        hipLaunchKernelGGL(k_root<double>, dim3(c->nsys), dim3(256), 0, st, bb, c->sys_blk0, c->sys_off,
                           c->sys_geom, c->dsc);
        if (levels == kLevelsTiny)
            hipLaunchKernelGGL((k_keys<kLevelsTiny, double>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x64,
                               c->y64, c->z64, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, nullptr, nullptr);
        else if (levels == kLevelsFast)
            hipLaunchKernelGGL((k_keys<kLevelsFast, double>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x64,
                               c->y64, c->z64, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, nullptr, nullptr);
        else if (levels == kLevelsShort)
            hipLaunchKernelGGL((k_keys<kLevelsShort, double>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x64,
                               c->y64, c->z64, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, nullptr, nullptr);
        else
            hipLaunchKernelGGL((k_keys<kLevelsDeep, double>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x64,
                               c->y64, c->z64, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, nullptr, nullptr);
        return;
    }
    hipLaunchKernelGGL(k_bbox<float>, dim3(c->n_blocks), dim3(kSortThreads), 0, st, c->x, c->y, c->z,
                       c->blk_start, c->blk_count, c->blk_bbox);
    hipLaunchKernelGGL(k_root<float>, dim3(c->nsys), dim3(256), 0, st, c->blk_bbox, c->sys_blk0, c->sys_off,
                       c->sys_geom, c->dsc);
    if (levels == kLevelsTiny)
        hipLaunchKernelGGL((k_keys<kLevelsTiny, float>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x, c->y,
                           c->z, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, c->m, c->stage4);
    else if (levels == kLevelsFast)
        hipLaunchKernelGGL((k_keys<kLevelsFast, float>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x, c->y,
                           c->z, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, c->m, c->stage4);
    else if (levels == kLevelsShort)
        hipLaunchKernelGGL((k_keys<kLevelsShort, float>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x, c->y,
                           c->z, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, c->m, c->stage4);
    else
        hipLaunchKernelGGL((k_keys<kLevelsDeep, float>), dim3(grid_for(c->n, 256)), dim3(256), 0, st, c->x, c->y,
                           c->z, c->pos_sys, c->sys_geom, c->n, c->key_a, c->klo_a, c->val_a, c->m, c->stage4);
}

// ============================================================================ radix sort
// LSD radix sort, 8-bit digits, segmented by system: the digit counters of a system's
// blocks form one contiguous region (digit-major, block-minor), so a single exclusive
// scan over all counters yields, for every (block, digit), the output offset inside its
// own system's range.  The scatter ranks keys stably inside the block with wave ballots
// and per-wave LDS digit counts.

__device__ __forceinline__ int hist_index(int b, int d, const int32_t* blk_sys, const int32_t* sys_blk0) {
    const int s = blk_sys[b];
    const int b0 = sys_blk0[s];
    const int nb = sys_blk0[s + 1] - b0;
    return kRadix * b0 + d * nb + (b - b0);
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t* __restrict__ keys, int shift,
                                                             const int32_t* __restrict__ blk_sys,
                                                             const int32_t* __restrict__ blk_start,
                                                             const int32_t* __restrict__ blk_count,
                                                             const int32_t* __restrict__ sys_blk0,
                                                             uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_cnt[kRadix];
    const int b = blockIdx.x;
    s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int start = blk_start[b], cnt = blk_count[b];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int k = r * kSortThreads + threadIdx.x;
        if (k < cnt) {
            const int d = static_cast<int>((keys[start + k] >> shift) & (kRadix - 1));
            atomicAdd(&s_cnt[d], 1u);
        }
    }
    __syncthreads();
    hist[hist_index(b, threadIdx.x, blk_sys, sys_blk0)] = s_cnt[threadIdx.x];
}

// Each wave owns a contiguous slice of the block's keys (kSortItems rounds of 64), so the
// block's input order is (wave, round, lane).  A wave ranks its keys stably by digit with
// ballots and a running count per digit in its own LDS row, with no block barrier; the
// per-wave counts then give each digit's offset per wave and the block-local digit offsets.
// The keys are placed in LDS in digit order and written out by consecutive threads to
// consecutive positions of each digit's output run, so the global writes come in runs
// instead of one scattered word per key.
template <bool EXTRA>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, const uint64_t* __restrict__ xin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t* __restrict__ xout, int shift,
    const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_start, const int32_t* __restrict__ blk_count,
    const int32_t* __restrict__ sys_blk0, const uint32_t* __restrict__ offs) {
    constexpr int kWaves = kSortThreads / 64;
    constexpr int kSlice = 64 * kSortItems;  // keys per wave
    static_assert(kRadix == kSortThreads, "one digit per thread");
    __shared__ uint32_t s_wrun[kWaves][kRadix];  // a wave's keys per digit so far, then its offset in the digit
    __shared__ uint32_t s_run[kRadix];           // block-local digit offsets
    __shared__ uint32_t s_gbase[kRadix];         // global output position of the digit's first key of this block
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ uint64_t s_key[kSortTile];
    __shared__ uint32_t s_val[kSortTile];
    __shared__ uint64_t s_xtra[EXTRA ? kSortTile : 1];
    const int b = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    s_gbase[threadIdx.x] = offs[hist_index(b, threadIdx.x, blk_sys, sys_blk0)];
    const int start = blk_start[b], cnt = blk_count[b];
    uint64_t key[kSortItems], xtra[kSortItems];
    uint32_t val[kSortItems], loc[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int k = wave * kSlice + r * 64 + lane;
        key[r] = 0;
        val[r] = 0;
        xtra[r] = 0;
        if (k < cnt) {
            key[r] = kin[start + k];
            val[r] = vin[start + k];
            if (EXTRA) xtra[r] = xin[start + k];
        }
    }
    uint32_t* wrun = s_wrun[wave];
#pragma unroll
    for (int j = 0; j < kRadix / 64; ++j) wrun[j * 64 + lane] = 0;
    wave_sync();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int k = wave * kSlice + r * 64 + lane;
        const bool valid = k < cnt;
        const int d = static_cast<int>((key[r] >> shift) & (kRadix - 1));
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int bit = 0; bit < kRadixBits; ++bit) {
            const bool set = (d >> bit) & 1;
            const uint64_t m = ballot(set);
            peers &= set ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        loc[r] = valid ? wrun[d] + static_cast<uint32_t>(rank) : 0u;  // rank among the wave's keys of digit d
        wave_sync();
        if (valid && rank == 0) wrun[d] += static_cast<uint32_t>(__popcll(peers));
        wave_sync();
    }
    __syncthreads();
    // digit t (one per thread): its per-wave counts become per-wave offsets inside the digit,
    // and its total enters the exclusive scan over digits
    {
        const int t = threadIdx.x;
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t cw = s_wrun[w][t];
            s_wrun[w][t] = c;
            c += cw;
        }
        uint32_t x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(x), o, 64));
            if (lane >= o) x += y;
        }
        if (lane == 63) s_wsum[wave] = x;
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < wave; ++w) before += s_wsum[w];
        s_run[t] = before + x - c;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int k = wave * kSlice + r * 64 + lane;
        if (k < cnt) {
            const int d = static_cast<int>((key[r] >> shift) & (kRadix - 1));
            const uint32_t p = s_run[d] + s_wrun[wave][d] + loc[r];
            s_key[p] = key[r];
            s_val[p] = val[r];
            if (EXTRA) s_xtra[p] = xtra[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int p = r * kSortThreads + threadIdx.x;
        if (p < cnt) {
            const uint64_t kk = s_key[p];
            const int d = static_cast<int>((kk >> shift) & (kRadix - 1));
            const uint32_t pos = s_gbase[d] + static_cast<uint32_t>(p) - s_run[d];
            kout[pos] = kk;
            vout[pos] = s_val[p];
            if (EXTRA) xout[pos] = s_xtra[p];
        }
    }
}

// After an even number of passes the sorted data is back in key_a/val_a(/klo_a), after an
// odd number in key_b/val_b.
void launch_sort(nbg_ctx* c, int levels) {
    hipStream_t st = c->stream;
    const int nb = c->n_blocks;
    const int64_t nh = static_cast<int64_t>(kRadix) * nb;
    auto pass = [&](uint64_t* kin, uint64_t* kout, uint32_t* vin, uint32_t* vout, uint64_t* xin, uint64_t* xout,
                    int shift) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, st, kin, shift, c->blk_sys, c->blk_start,
                           c->blk_count, c->sys_blk0, c->hist);
        excl_scan<uint32_t>(st, c->hist, c->hist, nh, reinterpret_cast<uint32_t*>(c->scan_part), nullptr);
        if (xin)
            hipLaunchKernelGGL(k_radix_scatter<true>, dim3(nb), dim3(kSortThreads), 0, st, kin, vin, xin, kout, vout,
                               xout, shift, c->blk_sys, c->blk_start, c->blk_count, c->sys_blk0, c->hist);
        else
            hipLaunchKernelGGL(k_radix_scatter<false>, dim3(nb), dim3(kSortThreads), 0, st, kin, vin, nullptr, kout,
                               vout, nullptr, shift, c->blk_sys, c->blk_start, c->blk_count, c->sys_blk0, c->hist);
    };
    if (levels == kLevelsDeep) {
        // low word first (levels 22..40), carrying the high word, then the high word
        for (int p = 0; p < 8; ++p) {
            const bool even = (p & 1) == 0;
            pass(even ? c->klo_a : c->klo_b, even ? c->klo_b : c->klo_a, even ? c->val_a : c->val_b,
                 even ? c->val_b : c->val_a, even ? c->key_a : c->key_b, even ? c->key_b : c->key_a, 8 * p);
        }
        for (int p = 0; p < 8; ++p) {
            const bool even = (p & 1) == 0;
            pass(even ? c->key_a : c->key_b, even ? c->key_b : c->key_a, even ? c->val_a : c->val_b,
                 even ? c->val_b : c->val_a, even ? c->klo_a : c->klo_b, even ? c->klo_b : c->klo_a, 8 * p);
        }
    } else {
        // the key's bits 3 (21 - levels) .. 62: 8 passes for 21 levels, 6 for 16 (bits 15..62),
        // 5 for 13 (bits 24..62; the sorted data then ends in the b buffers)
        const int low = 3 * (kLevelsShort - levels);
        const int passes = (63 - low + 7) / 8;
        for (int p = 0; p < passes; ++p) {
            const bool even = (p & 1) == 0;
            pass(even ? c->key_a : c->key_b, even ? c->key_b : c->key_a, even ? c->val_a : c->val_b,
                 even ? c->val_b : c->val_a, nullptr, nullptr, low + 8 * p);
        }
        if (passes & 1) {
            c->key_sorted = c->key_b;
            c->klo_sorted = c->klo_b;
            c->perm = c->val_b;
            return;
        }
    }
    c->key_sorted = c->key_a;
    c->klo_sorted = c->klo_a;
    c->perm = c->val_a;
}

// ============================================================================ topology

template <int LEVELS>
__device__ __forceinline__ int cpl_at(const uint64_t* __restrict__ kh, const uint64_t* __restrict__ kl, int i, int j) {
    // (a 16-level key holds no deeper levels: equal there is as far as it can tell)
    if (LEVELS <= kLevelsShort) return min(cpl_short(kh[i], kh[j]), LEVELS);
    return cpl_deep(kh[i], kl[i], kh[j], kl[j]);
}

// Number of particles in the depth-d cell that starts at sorted position i (its system
// ends at s1): the first j > i outside the cell, by bisection on the common prefix.
template <int LEVELS>
__device__ __forceinline__ int cell_count(const uint64_t* __restrict__ kh, const uint64_t* __restrict__ kl, int i,
                                          int s1, int d) {
    int lo = i + 1, hi = s1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cpl_at<LEVELS>(kh, kl, i, mid) >= d)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - i;
}
// Leaf depth and common prefix with the predecessor, per sorted position; the number of
// nodes and leaves each position starts, packed for one 64-bit scan.
template <int LEVELS>
__global__ __launch_bounds__(256) void k_leaf_depth(const uint64_t* __restrict__ kh, const uint64_t* __restrict__ kl,
                                                    const int32_t* __restrict__ pos_sys,
                                                    const int32_t* __restrict__ sys_off, int n,
                                                    int32_t* __restrict__ leaf_depth, int32_t* __restrict__ cpl_prev,
                                                    uint64_t* __restrict__ node_scan, DevScalars* __restrict__ sc,
                                                    int32_t* __restrict__ blk_depth, uint64_t* __restrict__ blk_sum) {
    __shared__ int s_dmax[4];
    __shared__ uint64_t s_sum[4];
    uint64_t nsv = 0;  // this position's node_scan word (0 past n)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int D = 0;  // positions past n report depth 0 to the block maximum
    if (i < n) {
        const int s = pos_sys[i];
        // the +-8 neighbours' keys are loaded (clamped to the array) while the system's
        // bounds are still in flight, which are only needed to mask the prefixes
        uint64_t hk[2 * kNcrit + 1], lk[LEVELS <= kLevelsShort ? 1 : 2 * kNcrit + 1];
#pragma unroll
        for (int j = 0; j <= 2 * kNcrit; ++j) {
            const int q = min(max(i + j - kNcrit, 0), n - 1);
            hk[j] = kh[q];
            if (LEVELS > kLevelsShort) lk[j] = kl[q];
        }
        auto cpl = [&](int a, int b) __attribute__((always_inline)) {
            if (LEVELS <= kLevelsShort) return min(cpl_short(hk[a], hk[b]), LEVELS);
            return cpl_deep(hk[a], lk[a], hk[b], lk[b]);
        };
        const int s0 = sys_off[s], s1 = sys_off[s + 1];
        int L[kNcrit], R[kNcrit];
#pragma unroll
        for (int k = 1; k <= kNcrit; ++k) {
            L[k - 1] = (i - k >= s0) ? cpl(kNcrit - k, kNcrit) : -1;
            R[k - 1] = (i + k < s1) ? cpl(kNcrit, kNcrit + k) : -1;
        }
        // cell size at depth d: 1 + #{L >= d} + #{R >= d}; the leaf depth is the first d with
        // size <= 8, i.e. one past v8, the 8th largest of the 16 prefix lengths.  L and R are
        // non-increasing (sorted keys), so v8 = max over a of min(L[a-1], R[7-a]) (taking a
        // from L and 8-a from R; an empty side counts as +infinity).
        int v8 = max(R[kNcrit - 1], L[kNcrit - 1]);
#pragma unroll
        for (int a = 1; a < kNcrit; ++a) v8 = max(v8, min(L[a - 1], R[kNcrit - 1 - a]));
        D = max(v8 + 1, 0);
        if (D > LEVELS) {
            if (LEVELS <= kLevelsShort) {
                atomicOr(&sc->overflow, 1);
                atomicOr(&sc->sticky_overflow, 1);
                D = LEVELS;
            } else {
                D = kMaxDepth;  // depth cap: the leaf keeps every particle (Tree.cpp:126)
            }
        }
        const int cp = L[0];  // -1 at the start of a system
        leaf_depth[i] = D;
        cpl_prev[i] = cp;
        const uint64_t ns = (D > cp) ? static_cast<uint64_t>(D - cp) : 0u;
        const uint64_t ls = (D > cp) ? 1u : 0u;
        nsv = (ns << 32) | ls;
        node_scan[i] = nsv;
    }
    // deepest leaf: one value per block (the totals kernel takes the maximum), no
    // same-address atomics; every thread reaches the one barrier
    // and the block's sum of the node_scan words (the scan's per-block partial)
    int dmax = D;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        dmax = max(dmax, __shfl_xor(dmax, o, 64));
        nsv += static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(nsv), o, 64));
    }
    if (lane_id() == 0) {
        s_dmax[threadIdx.x >> 6] = dmax;
        s_sum[threadIdx.x >> 6] = nsv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk_depth[blockIdx.x] = max(max(s_dmax[0], s_dmax[1]), max(s_dmax[2], s_dmax[3]));
        blk_sum[blockIdx.x] = (s_sum[0] + s_sum[1]) + (s_sum[2] + s_sum[3]);
    }
}

// Grand totals of the node scan and the deepest leaf -> device scalars (all threads of one
// block).
__device__ __forceinline__ void node_totals(const uint64_t* __restrict__ total, DevScalars* __restrict__ sc,
                                            int64_t cap_nodes, const int32_t* __restrict__ blk_depth, int nblk) {
    __shared__ int s_m[256];
    int m = 0;
    // eight loads in flight per thread and round (clamped indices repeat the last block)
    for (int b0 = 0; b0 < nblk; b0 += 8 * 256) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = blk_depth[min(b0 + u * 256 + static_cast<int>(threadIdx.x), nblk - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) m = max(m, v[u]);
    }
    s_m[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if (threadIdx.x < o) s_m[threadIdx.x] = max(s_m[threadIdx.x], s_m[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sc->max_depth = s_m[0];
        const uint64_t t = *total;
        const int64_t nodes = static_cast<int64_t>(t >> 32);
        sc->n_leaves = static_cast<int32_t>(t & 0xffffffffu);
        if (nodes > cap_nodes) {
            sc->cap_exceeded = 1;
            sc->sticky_cap = 1;
            sc->n_nodes = static_cast<int32_t>(cap_nodes);
        } else {
            sc->n_nodes = static_cast<int32_t>(nodes);
        }
    }
}

template <int LEVELS>
__global__ __launch_bounds__(256) void k_nodes(const uint64_t* __restrict__ kh, const uint64_t* __restrict__ kl,
                                               const int32_t* __restrict__ pos_sys,
                                               const int32_t* __restrict__ sys_off,
                                               const int32_t* __restrict__ leaf_depth,
                                               const int32_t* __restrict__ cpl_prev,
                                               const uint64_t* __restrict__ node_scan, int n, int64_t cap_nodes,
                                               int32_t* __restrict__ nd_start, int32_t* __restrict__ nd_count,
                                               int32_t* __restrict__ nd_depth, int32_t* __restrict__ nd_parent,
                                               int32_t* __restrict__ nd_oct, int32_t* __restrict__ nd_child,
                                               int32_t* __restrict__ leaf_list, int32_t* __restrict__ sys_root,
                                               DevScalars* __restrict__ sc, const int32_t* __restrict__ blk_depth,
                                               int nblk) {
    // block 0 also publishes the totals (a launch of its own cost a few microseconds; no
    // block of this kernel reads them)
    if (blockIdx.x == 0) node_totals(node_scan + n, sc, cap_nodes, blk_depth, nblk);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int D = leaf_depth[i];
    const int cp = cpl_prev[i];
    if (D <= cp) return;
    const int s = pos_sys[i];
    // the keys of positions i-8 .. i+8 (clamped to the array), all loads in flight together
    // while the system bounds are read: they give the leaf's count and, for 47% of the
    // chains of a 1M Plummer tree, the start of the parent cell, without a search (as
    // dependent loads behind the bounds tests they were 8 serial memory round trips)
    uint64_t hk[2 * kNcrit + 1], lk[LEVELS <= kLevelsShort ? 1 : 2 * kNcrit + 1];
#pragma unroll
    for (int j = 0; j <= 2 * kNcrit; ++j) {
        const int q = min(max(i + j - kNcrit, 0), n - 1);
        hk[j] = kh[q];
        if (LEVELS > kLevelsShort) lk[j] = kl[q];
    }
    auto cpl = [&](int a, int b) __attribute__((always_inline)) {
        if (LEVELS <= kLevelsShort) return min(cpl_short(hk[a], hk[b]), LEVELS);
        return cpl_deep(hk[a], lk[a], hk[b], lk[b]);
    };
    const int s0 = sys_off[s], s1 = sys_off[s + 1];
    const uint64_t sc_i = node_scan[i];
    const int64_t base = static_cast<int64_t>(sc_i >> 32);
    const int lbase = static_cast<int>(sc_i & 0xffffffffu);
    const uint64_t kl_i = (LEVELS <= kLevelsShort) ? 0u : lk[LEVELS <= kLevelsShort ? 0 : kNcrit];
    // particle count of the leaf (depth D): i and the following positions in its cell; a
    // leaf at the deepest level of the keys may hold more (two-word keys: the depth cap)
    int leaf_cnt = 1;
    if (D < LEVELS) {
#pragma unroll
        for (int k = 1; k <= kNcrit; ++k) leaf_cnt += (i + k < s1 && cpl(kNcrit, kNcrit + k) >= D);
    } else {
        leaf_cnt = cell_count<LEVELS>(kh, kl, i, s1, D);
    }
    // parent of the chain's first node (depth cp+1): the node of depth cp that holds i,
    // started by the first position of that cell, the first j in [s0, i] with cpl(j, i) >= cp
    // (non-decreasing in j)
    int parent0 = -1;
    if (cp >= 0) {
        int lo;
        if (i - kNcrit >= s0 && cpl(0, kNcrit) >= cp) {
            // further back than 8 (the particles of i's earlier siblings): gallop back from
            // i-8, then bisect the last interval (a bisection over the whole system range
            // took 20 dependent memory round trips)
            int in = i - kNcrit, out = s0 - 1;  // cpl(in, i) >= cp; out is outside (or before s0)
            for (int step = 2 * kNcrit;; step <<= 1) {
                const int j = i - step;
                if (j < s0) break;
                if (cpl_at<LEVELS>(kh, kl, j, i) < cp) {
                    out = j;
                    break;
                }
                in = j;
            }
            lo = out + 1;
            int hi = in;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cpl_at<LEVELS>(kh, kl, mid, i) >= cp)
                    hi = mid;
                else
                    lo = mid + 1;
            }
        } else {
            // within 8: i-1 .. i-k inside the cell, by the non-increasing prefixes
            int k = 0;
#pragma unroll
            for (int q = 1; q <= kNcrit; ++q) k += (i - q >= s0 && cpl(kNcrit - q, kNcrit) >= cp);
            lo = i - k;
        }
        parent0 = static_cast<int>(node_scan[lo] >> 32) + cp - (cpl_prev[lo] + 1);
    }
    for (int d = cp + 1; d <= D; ++d) {
        const int64_t id64 = base + (d - (cp + 1));
        if (id64 >= cap_nodes) {
            atomicOr(&sc->cap_exceeded, 1);
            atomicOr(&sc->sticky_cap, 1);
            return;
        }
        const int id = static_cast<int>(id64);
        // internal nodes' particle counts are summed by the upsweep
        if (d == D) nd_count[id] = leaf_cnt;
        int parent;
        if (d == 0) {
            parent = -1;
            sys_root[s] = id;
        } else {
            parent = (d > cp + 1) ? id - 1 : parent0;
        }
        nd_start[id] = i;
        nd_depth[id] = d;
        nd_parent[id] = parent;
        nd_oct[id] = (d > 0) ? key_digit(hk[kNcrit], kl_i, d) : 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) nd_child[8 * id + c] = -1;
        if (d == D) leaf_list[lbase] = id;
    }
}

__global__ __launch_bounds__(256) void k_link(const int32_t* __restrict__ nd_parent, const int32_t* __restrict__ nd_oct,
                                              const DevScalars* __restrict__ sc, int32_t* __restrict__ nd_child) {
    const int nn = sc->n_nodes;
    for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += gridDim.x * blockDim.x) {
        const int p = nd_parent[id];
        if (p >= 0) nd_child[8 * p + nd_oct[id]] = id;
    }
}

// Child counts, FP64 geometry (replaying Tree.cpp:96-99 along the key), leaf info word and
// each block's count of internal nodes per depth (counted in LDS, written as one row entry
// per depth: global same-address atomics from every block cost more than the rest).
template <int LEVELS>
__global__ __launch_bounds__(256) void k_node_finish(const uint64_t* __restrict__ kh, const uint64_t* __restrict__ kl,
                                                     const int32_t* __restrict__ pos_sys,
                                                     const int32_t* __restrict__ sys_off,
                                                     const double* __restrict__ sys_geom,
                                                     const int32_t* __restrict__ nd_start,
                                                     int32_t* __restrict__ nd_count,
                                                     const int32_t* __restrict__ nd_depth,
                                                     const int32_t* __restrict__ nd_child, double* __restrict__ nd_geom,
                                                     float4* __restrict__ nd_rec, const DevScalars* __restrict__ sc,
                                                     uint32_t* __restrict__ dep_hist) {
    __shared__ uint32_t s_hist[64];
    if (threadIdx.x < 64) s_hist[threadIdx.x] = 0;
    __syncthreads();
    // A truncated node pool (the host reports NBG_E_CAPACITY) can leave an internal node
    // without its children and a leaf without its count: write nothing from it.
    const int nn = sc->cap_exceeded ? 0 : sc->n_nodes;
    for (int base = blockIdx.x * blockDim.x; base < nn; base += gridDim.x * blockDim.x) {
        const int id = base + threadIdx.x;
        if (id >= nn) break;
        int nk = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) nk += nd_child[8 * id + c] >= 0;
        const int i = nd_start[id];
        const int d = nd_depth[id];
        const int s = pos_sys[i];
        const uint64_t h_ = kh[i];
        const uint64_t l_ = (LEVELS <= kLevelsShort) ? 0u : kl[i];
        double cx = sys_geom[4 * s + 0], cy = sys_geom[4 * s + 1], cz = sys_geom[4 * s + 2], h = sys_geom[4 * s + 3];
        for (int l = 1; l <= d; ++l) {
            const int oct = key_digit(h_, l_, l);
            h = 0.5 * h;
            cx = cx + ((oct & 1) ? h : -h);
            cy = cy + ((oct & 2) ? h : -h);
            cz = cz + ((oct & 4) ? h : -h);
        }
        nd_geom[4 * id + 0] = cx;
        nd_geom[4 * id + 1] = cy;
        nd_geom[4 * id + 2] = cz;
        nd_geom[4 * id + 3] = h;
        const double size = 2.0 * h;
        // an internal node's particle count is its children's sum, added by the upsweep
        const int cnt = (nk == 0) ? nd_count[id] : 0;
        const uint32_t info = (nk == 0) ? (kLeafBit | static_cast<uint32_t>(cnt)) : static_cast<uint32_t>(nk);
        const float4 aux = make_float4(0.0f, static_cast<float>(size * size), 0.0f, __uint_as_float(info));
        // the walk's record: the first particle and the aux word (the upsweep adds the rest)
        reinterpret_cast<int*>(nd_rec + 4 * id + 2)[2] = i;
        nd_rec[4 * id + 3] = aux;
        if (nk > 0) atomicAdd(&s_hist[d], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kInternalDepths<LEVELS>) dep_hist[threadIdx.x * kNodeBlocks + blockIdx.x] = s_hist[threadIdx.x];
}

// Internal nodes grouped by depth.  The per-block counts, laid out depth-major, are
// scanned exclusively (depth_ranges): entry (d, b) is then the first slot of block b's nodes
// of depth d, and row d's tile sum the number of nodes at depth d.  Each block ranks its
// nodes per depth in LDS and places them from those slots; it covers the same node ids as
// in k_node_finish (same grid, same stride).
__global__ __launch_bounds__(256) void k_depth_scatter(const int32_t* __restrict__ nd_depth,
                                                       const float4* __restrict__ nd_rec,
                                                       const DevScalars* __restrict__ sc,
                                                       const uint32_t* __restrict__ dep_off, int ndepth,
                                                       int32_t* __restrict__ depth_nodes) {
    __shared__ uint32_t s_cnt[64];
    __shared__ uint32_t s_base[64];
    const int nn = sc->cap_exceeded ? 0 : sc->n_nodes;
    if (threadIdx.x < ndepth) s_base[threadIdx.x] = dep_off[threadIdx.x * kNodeBlocks + blockIdx.x];
    for (int base = blockIdx.x * blockDim.x; base < nn; base += gridDim.x * blockDim.x) {
        if (threadIdx.x < 64) s_cnt[threadIdx.x] = 0;
        __syncthreads();
        const int id = base + threadIdx.x;
        int d = -1;
        uint32_t rank = 0;
        if (id < nn && !(__float_as_uint(nd_rec[4 * id + 3].w) & kLeafBit)) {
            d = nd_depth[id];
            rank = atomicAdd(&s_cnt[d], 1u);
        }
        __syncthreads();
        if (d >= 0) depth_nodes[s_base[d] + rank] = id;
        __syncthreads();
        if (threadIdx.x < ndepth) s_base[threadIdx.x] += s_cnt[threadIdx.x];
    }
}

// The input-order records packed (x, y, z, m; coalesced), so the per-leaf gather of the
// sorted particles reads one 16-byte record per particle instead of four scattered words.
__global__ __launch_bounds__(256) void k_pack4(const float* __restrict__ x, const float* __restrict__ y,
                                               const float* __restrict__ z, const float* __restrict__ m, int n,
                                               float4* __restrict__ pos4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pos4[i] = make_float4(x[i], y[i], z[i], m[i]);
}

template <int LEVELS>
void launch_build_t(nbg_ctx* c) {
    hipStream_t st = c->stream;
    const int n = c->n;
    const int g = grid_for(n, 256);
    const int gnodes = kNodeBlocks;  // grid-stride kernels over the device-side node count
    hipLaunchKernelGGL(k_leaf_depth<LEVELS>, dim3(g), dim3(256), 0, st, c->key_sorted, c->klo_sorted, c->pos_sys,
                       c->sys_off, n, c->leaf_depth, c->cpl_prev, c->node_scan, c->dsc, c->blk_depth, c->scan_part);
    // node_scan[n] receives the grand total (the array holds n+1 entries); the block sums of
    // k_leaf_depth (256 positions each) serve as the scan's partials
    static_assert(kScanTile % 256 == 0, "whole leaf-depth blocks per scan tile");
    const int tiles = grid_for(n, kScanTile);
    if (tiles <= kScanInline)
        hipLaunchKernelGGL(k_scan_down_inline<uint64_t>, dim3(tiles), dim3(kScanThreads), 0, st, c->node_scan,
                           c->node_scan, static_cast<int64_t>(n), c->scan_part, c->node_scan + n, kScanTile / 256);
    else
        excl_scan<uint64_t>(st, c->node_scan, c->node_scan, n, c->scan_part, c->node_scan + n);
    hipLaunchKernelGGL(k_nodes<LEVELS>, dim3(g), dim3(256), 0, st, c->key_sorted, c->klo_sorted, c->pos_sys,
                       c->sys_off, c->leaf_depth, c->cpl_prev, c->node_scan, n, c->cap_nodes, c->nd_start, c->nd_count,
                       c->nd_depth, c->nd_parent, c->nd_oct, c->nd_child, c->leaf_list, c->sys_root, c->dsc,
                       c->blk_depth, g);
    hipLaunchKernelGGL(k_link, dim3(gnodes), dim3(256), 0, st, c->nd_parent, c->nd_oct, c->dsc, c->nd_child);
    hipLaunchKernelGGL(k_node_finish<LEVELS>, dim3(gnodes), dim3(256), 0, st, c->key_sorted, c->klo_sorted,
                       c->pos_sys, c->sys_off, c->sys_geom, c->nd_start, c->nd_count, c->nd_depth, c->nd_child,
                       c->nd_geom, c->nd_rec, c->dsc, c->dep_hist);
    constexpr int nd = kInternalDepths<LEVELS>;
    excl_scan<uint32_t>(st, c->dep_hist, c->dep_off, static_cast<int64_t>(nd) * kNodeBlocks, c->dep_part, nullptr);
    hipLaunchKernelGGL(k_depth_scatter, dim3(gnodes), dim3(256), 0, st, c->nd_depth, c->nd_rec, c->dsc, c->dep_off,
                       nd, c->depth_nodes);
}

void launch_build(nbg_ctx* c, int levels) {
    if (levels == kLevelsTiny)
        launch_build_t<kLevelsTiny>(c);
    else if (levels == kLevelsFast)
        launch_build_t<kLevelsFast>(c);
    else if (levels == kLevelsShort)
        launch_build_t<kLevelsShort>(c);
    else
        launch_build_t<kLevelsDeep>(c);
}

// ============================================================================ upsweep
// Moments of every node (G-scaled masses): leaves from their particles, internal nodes
// from their children, deepest first (Tree.cpp:179-202, 205-254).  Centres of mass are
// reduced in FP64 (exact MAC decisions need them to FP64 accuracy; the walk's M2P uses
// their FP32 rounding); quadrupoles are accumulated in FP32 from offsets relative to the
// node's centre of mass.  A massless cell gets an undefined (NaN) centre of mass, as the
// incremental update of Tree.cpp:120-122 produces: the walk never uses it as a source, and
// a massless group then opens every cell, as in the reference.

// A node's FP32 record, one 64-byte line: [0] (com, G*M), [1] G*(Qxx, Qyy, Qxy, Qxz),
// [2] G*(Qyz, Qzz = -(Qxx + Qyy)), first sorted particle, -, [3] the aux word (G*Qyz, size^2,
// radius, info).  The walk's far field reads [0..2] (one scalar-cache line per cell instead
// of one from each of three arrays), its traversal [0], [3] and a leaf's first particle (one
// line per node).  k_node_finish writes the particle and [3].y, [3].w; the moments the rest.
__device__ __forceinline__ void rec_store(float4* __restrict__ nd_rec, int id, float4 com, float4 quad, float qyz,
                                          float rmax) {
    nd_rec[4 * id + 0] = com;
    nd_rec[4 * id + 1] = quad;
    reinterpret_cast<float2*>(nd_rec + 4 * id + 2)[0] = make_float2(qyz, -(quad.x + quad.y));
    float* aux = reinterpret_cast<float*>(nd_rec + 4 * id + 3);
    aux[0] = qyz;
    aux[2] = rmax;
}

// Per leaf, one lane each: its members put in ascending original index (the reference
// bucket order, so `order` lists them as Finalize does), their packed records gathered into
// sorted order as (x, y, z, G*m), the sorted-position -> leaf map, and the leaf's moments.
// F64: the positions of an FP64 state (x64, y64, z64, input order) give the FP64 centre
// of mass and radius, and the quadrupole's offsets from that centre (rounded to FP32);
// otherwise the FP32 records do.
template <bool QUAD, bool MOM, bool F64>
__global__ __launch_bounds__(256) void k_leaves(const int32_t* __restrict__ leaf_list,
                                                const int32_t* __restrict__ nd_start,
                                                const int32_t* __restrict__ nd_count,
                                                const float4* __restrict__ pos4, uint32_t* __restrict__ perm, float Gf,
                                                double G, const DevScalars* __restrict__ sc,
                                                float4* __restrict__ spos, int32_t* __restrict__ leaf_of,
                                                float4* __restrict__ nd_rec,
                                                double4* __restrict__ nd_com64, double* __restrict__ nd_rmax64,
                                                float4* __restrict__ nd_mom, const double* __restrict__ x64,
                                                const double* __restrict__ y64, const double* __restrict__ z64) {
    const int nl = sc->cap_exceeded ? 0 : sc->n_leaves;  // a truncated pool: nothing to do
    for (int li = blockIdx.x * blockDim.x + threadIdx.x; li < nl; li += gridDim.x * blockDim.x) {
        const int id = leaf_list[li];
        const int st = nd_start[id], cnt = nd_count[id];
        // A leaf below the depth cap holds at most kNcrit particles: they are sorted and
        // gathered in registers (all loads in flight together) and the passes below run on
        // them; a capped leaf of any size is sorted in place and streamed from memory in
        // each pass.  Both forms accumulate in ascending position order.
        const bool small = cnt <= kNcrit;
        float4 pr[kNcrit];
        double wr[kNcrit];
        double3 pd[F64 ? kNcrit : 1];
        if (small) {
            // loads from clamped positions, all in flight together (loads behind the
            // k < cnt tests were issued one round trip after another)
            uint32_t v[kNcrit];
#pragma unroll
            for (int k = 0; k < kNcrit; ++k) {
                const uint32_t pv = perm[st + min(k, cnt - 1)];
                v[k] = (k < cnt) ? pv : 0xffffffffu;
            }
            // insertion network over <= 8 values (absent ones sort last)
#pragma unroll
            for (int a = 1; a < kNcrit; ++a) {
#pragma unroll
                for (int b = a; b > 0; --b) {
                    const uint32_t lo = min(v[b - 1], v[b]), hi = max(v[b - 1], v[b]);
                    v[b - 1] = lo;
                    v[b] = hi;
                }
            }
#pragma unroll
            for (int u = 0; u < kNcrit; ++u) {
                const uint32_t q = (u < cnt) ? v[u] : v[0];
                const float4 r = pos4[q];
                pr[u] = make_float4(r.x, r.y, r.z, Gf * r.w);
                wr[u] = static_cast<double>(r.w);
                if (F64) pd[u] = make_double3(x64[q], y64[q], z64[q]);
            }
#pragma unroll
            for (int u = 0; u < kNcrit; ++u) {
                if (u < cnt) {
                    perm[st + u] = v[u];
                    spos[st + u] = pr[u];
                    leaf_of[st + u] = id;
                }
            }
        } else {
            // depth-capped leaf: plain insertion sort in place, then the gather
            for (int a = st + 1; a < st + cnt; ++a) {
                const uint32_t key = perm[a];
                int b = a - 1;
                while (b >= st && perm[b] > key) {
                    perm[b + 1] = perm[b];
                    --b;
                }
                perm[b + 1] = key;
            }
            for (int k = st; k < st + cnt; ++k) {
                const float4 r = pos4[perm[k]];
                spos[k] = make_float4(r.x, r.y, r.z, Gf * r.w);
                leaf_of[k] = id;
            }
        }
        // calls f(p, w, position in FP64) for every member in ascending order
        auto each = [&](auto f) __attribute__((always_inline)) {
            if (small) {
#pragma unroll
                for (int u = 0; u < kNcrit; ++u)
                    if (u < cnt) f(pr[u], wr[u], F64 ? pd[u] : make_double3(pr[u].x, pr[u].y, pr[u].z));
            } else {
                for (int k = st; k < st + cnt; ++k) {
                    const uint32_t i = perm[k];
                    const float4 r = pos4[i];
                    const float4 p = make_float4(r.x, r.y, r.z, Gf * r.w);
                    f(p, static_cast<double>(r.w), F64 ? make_double3(x64[i], y64[i], z64[i])
                                                       : make_double3(p.x, p.y, p.z));
                }
            }
        };
        double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
        each([&](const float4&, double w, const double3& q) __attribute__((always_inline)) {
            M += w;
            sx += w * q.x;
            sy += w * q.y;
            sz += w * q.z;
        });
        double gx = __longlong_as_double(0x7ff8000000000000ll), gy = gx, gz = gx;
        if (M != 0.0) {
            gx = sx / M;
            gy = sy / M;
            gz = sz / M;
        }
        const float fx = static_cast<float>(gx), fy = static_cast<float>(gy), fz = static_cast<float>(gz);
        float qxx = 0.f, qyy = 0.f, qxy = 0.f, qxz = 0.f, qyz = 0.f;
        double rmax = 0.0;
        if (M > 0.0) {
            each([&](const float4& p, double, const double3& q) __attribute__((always_inline)) {
                const double ex = q.x - gx, ey = q.y - gy, ez = q.z - gz;
                rmax = fmax(rmax, sqrt(ex * ex + ey * ey + ez * ez));
                if (QUAD && cnt > 1) {
                    const float dx = F64 ? static_cast<float>(ex) : p.x - fx;
                    const float dy = F64 ? static_cast<float>(ey) : p.y - fy;
                    const float dz = F64 ? static_cast<float>(ez) : p.z - fz;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    qxx += p.w * (3.f * dx * dx - d2);
                    qyy += p.w * (3.f * dy * dy - d2);
                    qxy += p.w * 3.f * dx * dy;
                    qxz += p.w * 3.f * dx * dz;
                    qyz += p.w * 3.f * dy * dz;
                }
            });
        }
        if (MOM) {
            // second and third moments about the FP32 com, G*m weights (learned correction):
            // [xx yy zz xy] [xz yz xxx yyy] [zzz xxy xxz xyy] [yyz xzz yzz xyz]
            float mo[16] = {};
            if (M > 0.0) {
                each([&](const float4& p, double, const double3&) __attribute__((always_inline)) {
                    const float dx = p.x - fx, dy = p.y - fy, dz = p.z - fz, w = p.w;
                    mo[0] += w * dx * dx;
                    mo[1] += w * dy * dy;
                    mo[2] += w * dz * dz;
                    mo[3] += w * dx * dy;
                    mo[4] += w * dx * dz;
                    mo[5] += w * dy * dz;
                    mo[6] += w * dx * dx * dx;
                    mo[7] += w * dy * dy * dy;
                    mo[8] += w * dz * dz * dz;
                    mo[9] += w * dx * dx * dy;
                    mo[10] += w * dx * dx * dz;
                    mo[11] += w * dx * dy * dy;
                    mo[12] += w * dy * dy * dz;
                    mo[13] += w * dx * dz * dz;
                    mo[14] += w * dy * dz * dz;
                    mo[15] += w * dx * dy * dz;
                });
            }
            nd_mom[4 * id + 0] = make_float4(mo[0], mo[1], mo[2], mo[3]);
            nd_mom[4 * id + 1] = make_float4(mo[4], mo[5], mo[6], mo[7]);
            nd_mom[4 * id + 2] = make_float4(mo[8], mo[9], mo[10], mo[11]);
            nd_mom[4 * id + 3] = make_float4(mo[12], mo[13], mo[14], mo[15]);
        }
        nd_com64[id] = make_double4(gx, gy, gz, M);
        nd_rmax64[id] = rmax;
        const float4 com = make_float4(fx, fy, fz, static_cast<float>(G * M));
        const float4 quad = make_float4(qxx, qyy, qxy, qxz);
        rec_store(nd_rec, id, com, quad, qyz, static_cast<float>(rmax));
    }
}

// An empty asm that takes v as input and output: the compiler must have the loaded value
// in registers here and cannot move its load into a later branch.
__device__ __forceinline__ void pin(float4& v) {
    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
}

// What a node's parent combines from it: FP64 centre of mass and mass, the aux word (x:
// qyz, z: radius), the quadrupole's other four terms and the particle count.
struct UpRec {
    double4 cd;
    float4 ca;
    float4 cq;
    int cn;
};

// One internal node from its eight child slots (ch[c] < 0: no child; its slot may hold
// anything).  rec(c) gives child c's record, mom(c, j) word j of its learned moments.
// Branch-free over the children: every child's record is fetched before the first is used
// (behind per-child branches they were 40-odd dependent memory round trips per node, 7 us
// for the root alone), and every accumulation an absent child would make is discarded by a
// select that keeps the arithmetic of the skip-by-branch form term for term.
// F64: the offsets of the children's centres of mass (quadrupole shifts and radius) are
// taken in FP64 before rounding (an FP64 state); otherwise from their FP32 roundings.
template <bool QUAD, bool MOM, bool F64, class Rec, class Mom>
__device__ __forceinline__ void up_combine(const int (&ch)[8], double G, Rec rec, Mom mom, UpRec& out,
                                           float (&mo)[16]) {
    static_assert(!(MOM && F64), "the learned moments are built on FP32 states only");
    double4 cd[8];
    float4 cc[8], ca[8], cq[8];
    int cn[8];
    // the learned moments of the first four children are fetched with the records (they
    // are used last; fetched after the centre of mass they added two memory round trips)
    float4 cm[2][16];
#pragma unroll
    for (int c = 0; c < (MOM ? 4 : 0); ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) cm[0][4 * c + j] = mom(c, j);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const UpRec r = rec(c);
        cd[c] = r.cd;
        ca[c] = r.ca;
        cq[c] = r.cq;
        cn[c] = r.cn;
    }
    // Values used only on one side of a select would otherwise be loaded inside a branch
    // (one memory round trip per child); pinning them here keeps one wait.
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        // a child's FP32 record is the rounding of its FP64 one (both written from the same
        // values), so it is rebuilt instead of gathered
        cc[c] = make_float4(static_cast<float>(cd[c].x), static_cast<float>(cd[c].y), static_cast<float>(cd[c].z),
                            static_cast<float>(G * cd[c].w));
        pin(ca[c]);
        if (QUAD) pin(cq[c]);
    }
    double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const double w = cd[c].w;
        const bool use = ch[c] >= 0 && w != 0.0;
        M = use ? M + w : M;
        sx = use ? sx + w * cd[c].x : sx;
        sy = use ? sy + w * cd[c].y : sy;
        sz = use ? sz + w * cd[c].z : sz;
    }
    double gx = __longlong_as_double(0x7ff8000000000000ll), gy = gx, gz = gx;
    if (M != 0.0) {
        gx = sx / M;
        gy = sy / M;
        gz = sz / M;
    }
    const float fx = static_cast<float>(gx), fy = static_cast<float>(gy), fz = static_cast<float>(gz);
    float qxx = 0.f, qyy = 0.f, qxy = 0.f, qxz = 0.f, qyz = 0.f, rmax = 0.f;
    const bool massive = M > 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const bool live = massive && ch[c] >= 0;
        const float dx = F64 ? static_cast<float>(cd[c].x - gx) : cc[c].x - fx;
        const float dy = F64 ? static_cast<float>(cd[c].y - gy) : cc[c].y - fy;
        const float dz = F64 ? static_cast<float>(cd[c].z - gz) : cc[c].z - fz;
        const float d2 = dx * dx + dy * dy + dz * dz;
        rmax = live ? fmaxf(rmax, ca[c].z + sqrtf(d2)) : rmax;
        const bool heavy = live && cc[c].w > 0.f;
        if (QUAD) {
            const float w = cc[c].w;
            qxx = heavy ? qxx + (cq[c].x + w * (3.f * dx * dx - d2)) : qxx;
            qyy = heavy ? qyy + (cq[c].y + w * (3.f * dy * dy - d2)) : qyy;
            qxy = heavy ? qxy + (cq[c].z + w * 3.f * dx * dy) : qxy;
            qxz = heavy ? qxz + (cq[c].w + w * 3.f * dx * dz) : qxz;
            qyz = heavy ? qyz + (ca[c].x + w * 3.f * dy * dz) : qyz;
        }
    }
    // the learned moments, four children at a time (all eight held at once took ~300
    // registers): the last four children's words are fetched before the first four's are
    // used; each moment still sums the children in order
#pragma unroll
    for (int i = 0; i < 16; ++i) mo[i] = 0.f;
#pragma unroll
    for (int c = 0; c < (MOM ? 4 : 0); ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) cm[1][4 * c + j] = mom(4 + c, j);
#pragma unroll
    for (int h = 0; h < (MOM ? 2 : 0); ++h) {
        const int c0 = 4 * h;
#pragma unroll
        for (int c = 0; c < 4; ++c) pin(cm[h][4 * c]), pin(cm[h][4 * c + 1]), pin(cm[h][4 * c + 2]), pin(cm[h][4 * c + 3]);
#pragma unroll
        for (int cc4 = 0; cc4 < 4; ++cc4) {
            const int c = c0 + cc4;
            const bool heavy = massive && ch[c] >= 0 && cc[c].w > 0.f;
            const float dx = cc[c].x - fx, dy = cc[c].y - fy, dz = cc[c].z - fz;
            // parallel-axis shift of the child's moments to this com
            const float4 q0 = cm[h][4 * cc4 + 0], q1 = cm[h][4 * cc4 + 1], q2 = cm[h][4 * cc4 + 2],
                         q3 = cm[h][4 * cc4 + 3];
            const float m = cc[c].w;
            const float sxx = q0.x, syy = q0.y, szz = q0.z, sxy = q0.w, sxz = q1.x, syz = q1.y;
            float t[16];
            t[0] = sxx + m * dx * dx;
            t[1] = syy + m * dy * dy;
            t[2] = szz + m * dz * dz;
            t[3] = sxy + m * dx * dy;
            t[4] = sxz + m * dx * dz;
            t[5] = syz + m * dy * dz;
            t[6] = q1.z + 3.f * sxx * dx + m * dx * dx * dx;
            t[7] = q1.w + 3.f * syy * dy + m * dy * dy * dy;
            t[8] = q2.x + 3.f * szz * dz + m * dz * dz * dz;
            t[9] = q2.y + sxx * dy + 2.f * sxy * dx + m * dx * dx * dy;
            t[10] = q2.z + sxx * dz + 2.f * sxz * dx + m * dx * dx * dz;
            t[11] = q2.w + syy * dx + 2.f * sxy * dy + m * dx * dy * dy;
            t[12] = q3.x + syy * dz + 2.f * syz * dy + m * dy * dy * dz;
            t[13] = q3.y + szz * dx + 2.f * sxz * dz + m * dx * dz * dz;
            t[14] = q3.z + szz * dy + 2.f * syz * dz + m * dy * dz * dz;
            t[15] = q3.w + sxy * dz + sxz * dy + syz * dx + m * dx * dy * dz;
#pragma unroll
            for (int i = 0; i < 16; ++i) mo[i] = heavy ? mo[i] + t[i] : mo[i];
        }
    }
    int count = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) count += ch[c] >= 0 ? cn[c] : 0;
    out.cd = make_double4(gx, gy, gz, M);
    out.ca = make_float4(qyz, 0.f, rmax, 0.f);
    out.cq = make_float4(qxx, qyy, qxy, qxz);
    out.cn = count;
}

// An internal node's combined record to its arrays (the aux word's y and w stay as built).
template <bool MOM>
__device__ __forceinline__ void up_store(int id, const UpRec& r, const float (&mo)[16], double G,
                                         int32_t* __restrict__ nd_count, float4* __restrict__ nd_rec, double4* __restrict__ nd_com64,
                                         float4* __restrict__ nd_mom) {
    if (MOM) {
        nd_mom[4 * id + 0] = make_float4(mo[0], mo[1], mo[2], mo[3]);
        nd_mom[4 * id + 1] = make_float4(mo[4], mo[5], mo[6], mo[7]);
        nd_mom[4 * id + 2] = make_float4(mo[8], mo[9], mo[10], mo[11]);
        nd_mom[4 * id + 3] = make_float4(mo[12], mo[13], mo[14], mo[15]);
    }
    nd_count[id] = r.cn;
    nd_com64[id] = r.cd;
    const float4 com = make_float4(static_cast<float>(r.cd.x), static_cast<float>(r.cd.y), static_cast<float>(r.cd.z),
                                   static_cast<float>(G * r.cd.w));
    rec_store(nd_rec, id, com, r.cq, r.ca.x, r.ca.z);
}

// A node's record as its arrays hold it (a leaf's, or an internal node combined earlier).
template <bool QUAD>
__device__ __forceinline__ UpRec up_load(int i, const int32_t* __restrict__ nd_count,
                                         const float4* __restrict__ nd_rec,
                                         const double4* __restrict__ nd_com64) {
    UpRec r;
    r.cd = nd_com64[i];
    r.ca = nd_rec[4 * i + 3];
    if (QUAD) r.cq = nd_rec[4 * i + 1];
    r.cn = nd_count[i];
    return r;
}

// Combines internal node `id` from its children's records in memory.
template <bool QUAD, bool MOM, bool F64>
__device__ __forceinline__ void up_node_mem(int id, const int (&ch)[8], double G, int32_t* __restrict__ nd_count,
                                            float4* __restrict__ nd_rec,
                                            double4* __restrict__ nd_com64, float4* __restrict__ nd_mom, UpRec& out,
                                            float (&mo)[16]) {
    int src[8];  // an absent child reads the node's own slot (discarded)
#pragma unroll
    for (int c = 0; c < 8; ++c) src[c] = ch[c] >= 0 ? ch[c] : id;
    up_combine<QUAD, MOM, F64>(
        ch, G,
        [&](int c) __attribute__((always_inline)) { return up_load<QUAD>(src[c], nd_count, nd_rec, nd_com64); },
        [&](int c, int j) __attribute__((always_inline)) { return nd_mom[4 * src[c] + j]; }, out, mo);
    up_store<MOM>(id, out, mo, G, nd_count, nd_rec, nd_com64, nd_mom);
}

// One depth: a lane per internal node of that depth.
template <bool QUAD, bool MOM, bool F64>
__global__ __launch_bounds__(256) void k_up_internal(int depth, const int32_t* __restrict__ depth_nodes,
                                                     const int32_t* __restrict__ nd_child,
                                                     int32_t* __restrict__ nd_count, double G,
                                                     const uint32_t* __restrict__ dep_part,
                                                     const uint32_t* __restrict__ dep_off,
                                                     float4* __restrict__ nd_rec,
                                                     double4* __restrict__ nd_com64, float4* __restrict__ nd_mom) {
    const int cnt = static_cast<int>(dep_part[depth]);
    const int off = static_cast<int>(dep_off[depth * kNodeBlocks]);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int id = depth_nodes[off + k];
        const int4 c0 = reinterpret_cast<const int4*>(nd_child)[2 * id];
        const int4 c1 = reinterpret_cast<const int4*>(nd_child)[2 * id + 1];
        const int ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        UpRec r;
        float mo[16];
        up_node_mem<QUAD, MOM, F64>(id, ch, G, nd_count, nd_rec, nd_com64, nd_mom, r, mo);
    }
}

// A node's record into LDS slot t for its parent's combine: its own record, loaded with its
// child list (final for a leaf, unused for an internal node; loaded in a leaf branch it cost a
// round trip after the other lanes' combines in a wave that has both), or, for an internal
// node, combined from its children's records in memory (and written).
template <bool QUAD, bool MOM, bool F64>
__device__ __forceinline__ void up_child_to_lds(int c, int t, double G, const int32_t* __restrict__ nd_child,
                                                int32_t* __restrict__ nd_count, float4* __restrict__ nd_rec, double4* __restrict__ nd_com64,
                                                float4* __restrict__ nd_mom, UpRec* s_rec, float4* s_mom) {
    const int4 c0 = reinterpret_cast<const int4*>(nd_child)[2 * c];
    const int4 c1 = reinterpret_cast<const int4*>(nd_child)[2 * c + 1];
    UpRec r = up_load<QUAD>(c, nd_count, nd_rec, nd_com64);
    float mo[16];
    if (MOM) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 v = nd_mom[4 * c + j];
            mo[4 * j + 0] = v.x;
            mo[4 * j + 1] = v.y;
            mo[4 * j + 2] = v.z;
            mo[4 * j + 3] = v.w;
        }
    }
    const int gch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    const bool internal = max(max(max(c0.x, c0.y), max(c0.z, c0.w)), max(max(c1.x, c1.y), max(c1.z, c1.w))) >= 0;
    if (internal)
        up_node_mem<QUAD, MOM, F64>(c, gch, G, nd_count, nd_rec, nd_com64, nd_mom, r, mo);
    s_rec[t] = r;
    if (MOM) {
#pragma unroll
        for (int j = 0; j < 4; ++j) s_mom[4 * t + j] = make_float4(mo[4 * j], mo[4 * j + 1], mo[4 * j + 2], mo[4 * j + 3]);
    }
}

// Combines node `id` from the eight LDS slots g0..g0+7 (its children, ch[j] < 0: absent) and
// writes it (and, if t >= 0, its record into LDS slot t).
template <bool QUAD, bool MOM, bool F64>
__device__ __forceinline__ void up_node_lds(int id, const int (&ch)[8], int g0, int t, double G,
                                            int32_t* __restrict__ nd_count, float4* __restrict__ nd_rec, double4* __restrict__ nd_com64,
                                            float4* __restrict__ nd_mom, UpRec* s_rec, float4* s_mom) {
    UpRec r;
    float mo[16];
    up_combine<QUAD, MOM, F64>(
        ch, G, [&](int j) __attribute__((always_inline)) { return s_rec[g0 + j]; },
        [&](int j, int w) __attribute__((always_inline)) { return s_mom[4 * (g0 + j) + w]; }, r, mo);
    up_store<MOM>(id, r, mo, G, nd_count, nd_rec, nd_com64, nd_mom);
    if (t >= 0) {
        s_rec[t] = r;
        if (MOM) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                s_mom[4 * t + j] = make_float4(mo[4 * j], mo[4 * j + 1], mo[4 * j + 2], mo[4 * j + 3]);
        }
    }
}

// Two depths per launch, which halves the launches and the memory round trips per depth
// (index list, child lists) that bound this small, latency-bound work: 8 lanes per internal
// node of `depth`, lane a on its child a (at depth + 1).  An internal child is combined from
// its own children (done by the previous launch) and written; a leaf child's record is read.
// The 8 records go through LDS to the group's first lane, which combines the node.  Every
// node is combined by up_combine, as in the one-depth kernel, so the results are the same
// bits.
constexpr int kUpPairNodes = 32;  // nodes per 256-thread block
template <bool QUAD, bool MOM, bool F64>
__global__ __launch_bounds__(256) void k_up_pair(int depth, const int32_t* __restrict__ depth_nodes,
                                                 const int32_t* __restrict__ nd_child,
                                                 int32_t* __restrict__ nd_count, double G,
                                                 const uint32_t* __restrict__ dep_part,
                                                 const uint32_t* __restrict__ dep_off, float4* __restrict__ nd_rec, double4* __restrict__ nd_com64,
                                                 float4* __restrict__ nd_mom) {
    __shared__ UpRec s_rec[256];
    __shared__ float4 s_mom[MOM ? 4 * 256 : 1];
    __shared__ int s_ch[256];
    const int cnt = static_cast<int>(dep_part[depth]);
    const int off = static_cast<int>(dep_off[depth * kNodeBlocks]);
    const int t = threadIdx.x;
    const int a = t & 7, g0 = t & ~7;
    // block-uniform loop: every lane of a wave reaches each wave_sync
    for (int base = blockIdx.x * kUpPairNodes; base < cnt; base += gridDim.x * kUpPairNodes) {
        const int k = base + (t >> 3);
        const int id = (k < cnt) ? depth_nodes[off + k] : -1;
        const int c = (id >= 0) ? nd_child[8 * id + a] : -1;
        s_ch[t] = c;
        if (c >= 0)
            up_child_to_lds<QUAD, MOM, F64>(c, t, G, nd_child, nd_count, nd_rec, nd_com64,
                                            nd_mom, s_rec, s_mom);
        wave_sync();
        if (a == 0 && id >= 0) {
            int ch[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) ch[j] = s_ch[g0 + j];
            up_node_lds<QUAD, MOM, F64>(id, ch, g0, -1, G, nd_count, nd_rec, nd_com64, nd_mom,
                                        s_rec, s_mom);
        }
        wave_sync();
    }
}

// The top three depths (0, 1, 2) in one launch: one wave per internal node of depth 0 (a
// system root), lane 8a + b on grandchild b of child a (depth 2; an internal one is combined
// from its children, done by the previous launch); lane 8a then combines child a from its
// eight grandchildren's LDS records (or reads it, a leaf), and lane 0 the root from its
// children's.  `depth` is 0.
template <bool QUAD, bool MOM, bool F64>
__global__ __launch_bounds__(256) void k_up_top(int depth, const int32_t* __restrict__ depth_nodes,
                                                const int32_t* __restrict__ nd_child,
                                                int32_t* __restrict__ nd_count, double G,
                                                const uint32_t* __restrict__ dep_part,
                                                const uint32_t* __restrict__ dep_off, float4* __restrict__ nd_rec, double4* __restrict__ nd_com64,
                                                float4* __restrict__ nd_mom) {
    __shared__ UpRec s_rec[256];
    __shared__ float4 s_mom[MOM ? 4 * 256 : 1];
    __shared__ int s_ch[256];
    const int cnt = static_cast<int>(dep_part[depth]);
    const int off = static_cast<int>(dep_off[depth * kNodeBlocks]);
    const int t = threadIdx.x;
    const int lane = lane_id(), w0 = t - lane;  // the wave's first slot
    const int a = lane >> 3, b = lane & 7;
    // block-uniform loop: every lane of a wave reaches each wave_sync
    for (int base = blockIdx.x * (256 / 64); base < cnt; base += gridDim.x * (256 / 64)) {
        const int k = base + (t >> 6);
        const int id = (k < cnt) ? depth_nodes[off + k] : -1;
        const int ca = (id >= 0) ? nd_child[8 * id + a] : -1;
        const int g = (ca >= 0) ? nd_child[8 * ca + b] : -1;
        s_ch[t] = g;
        if (g >= 0)
            up_child_to_lds<QUAD, MOM, F64>(g, t, G, nd_child, nd_count, nd_rec, nd_com64,
                                            nd_mom, s_rec, s_mom);
        wave_sync();
        int ch[8];
        if (b == 0 && ca >= 0) {
            bool internal = false;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                ch[j] = s_ch[t + j];
                internal |= ch[j] >= 0;
            }
            if (internal) {
                up_node_lds<QUAD, MOM, F64>(ca, ch, t, t, G, nd_count, nd_rec, nd_com64,
                                            nd_mom, s_rec, s_mom);
            } else {
                s_rec[t] = up_load<QUAD>(ca, nd_count, nd_rec, nd_com64);
                if (MOM) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) s_mom[4 * t + j] = nd_mom[4 * ca + j];
                }
            }
        }
        wave_sync();
        if (b == 0) s_ch[t] = ca;  // (every lane read its group's slots before the sync)
        wave_sync();
        if (lane == 0 && id >= 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) ch[j] = s_ch[w0 + 8 * j];
            UpRec r;
            float mo[16];
            up_combine<QUAD, MOM, F64>(
                ch, G, [&](int j) __attribute__((always_inline)) { return s_rec[w0 + 8 * j]; },
                [&](int j, int wd) __attribute__((always_inline)) { return s_mom[4 * (w0 + 8 * j) + wd]; }, r, mo);
            up_store<MOM>(id, r, mo, G, nd_count, nd_rec, nd_com64, nd_mom);
        }
        wave_sync();
    }
}

void launch_upsweep(nbg_ctx* c, int levels, int multipole, double G, bool moments) {
    hipStream_t st = c->stream;
    const int gnodes = kNodeBlocks;
    const bool quad = multipole == NBG_QUADRUPOLE;
    const bool mom = moments && c->nd_mom != nullptr;
    const float Gf = static_cast<float>(G);
#define NBG_LEAVES(Q, M, F)                                                                                          \
    hipLaunchKernelGGL((k_leaves<Q, M, F>), dim3(gnodes), dim3(256), 0, st, c->leaf_list, c->nd_start, c->nd_count,   \
                       c->stage4, c->perm, Gf, G, c->dsc, c->spos, c->leaf_of, c->nd_rec, c->nd_com64,       \
                       c->nd_rmax64, c->nd_mom, c->x64, c->y64, c->z64)
    if (mom)
        NBG_LEAVES(true, true, false);
    else if (quad && c->f64_state)
        NBG_LEAVES(true, false, true);
    else if (quad)
        NBG_LEAVES(true, false, false);
    else if (c->f64_state)
        NBG_LEAVES(false, false, true);
    else
        NBG_LEAVES(false, false, false);
#undef NBG_LEAVES
    // internal nodes sit at depths 0..top (depth-40 cells are always leaves): two depths
    // per launch from the deepest, (top-1, top), (top-3, top-2), ..., and the top three in one
#define NBG_UP(K, D, GD)                                                                                             \
    do {                                                                                                             \
        if (mom)                                                                                                     \
            hipLaunchKernelGGL((K<true, true, false>), dim3(GD), dim3(256), 0, st, D, c->depth_nodes, c->nd_child,    \
                               c->nd_count, G, c->dep_part, c->dep_off, c->nd_rec, c->nd_com64, c->nd_mom);   \
        else if (quad && c->f64_state)                                                                               \
            hipLaunchKernelGGL((K<true, false, true>), dim3(GD), dim3(256), 0, st, D, c->depth_nodes, c->nd_child,    \
                               c->nd_count, G, c->dep_part, c->dep_off, c->nd_rec, c->nd_com64, c->nd_mom);   \
        else if (quad)                                                                                               \
            hipLaunchKernelGGL((K<true, false, false>), dim3(GD), dim3(256), 0, st, D, c->depth_nodes, c->nd_child,   \
                               c->nd_count, G, c->dep_part, c->dep_off, c->nd_rec, c->nd_com64, c->nd_mom);   \
        else if (c->f64_state)                                                                                       \
            hipLaunchKernelGGL((K<false, false, true>), dim3(GD), dim3(256), 0, st, D, c->depth_nodes, c->nd_child,   \
                               c->nd_count, G, c->dep_part, c->dep_off, c->nd_rec, c->nd_com64, c->nd_mom);   \
        else                                                                                                         \
            hipLaunchKernelGGL((K<false, false, false>), dim3(GD), dim3(256), 0, st, D, c->depth_nodes, c->nd_child,  \
                               c->nd_count, G, c->dep_part, c->dep_off, c->nd_rec, c->nd_com64, c->nd_mom);   \
    } while (0)
    const int top = (levels <= kLevelsShort) ? levels - 1 : kMaxDepth - 1;
    int d = top;
    if (top >= 2) {
        // pairs from the deepest down to depth 3, depth 3 alone if left over, then the top
        // three depths in one launch
        for (; d >= 4; d -= 2) NBG_UP(k_up_pair, d - 1, d - 1 < 5 ? 64 : 1024);
        if (d == 3) NBG_UP(k_up_internal, 3, 64);
        NBG_UP(k_up_top, 0, 64);
    } else {
        for (; d >= 1; d -= 2) NBG_UP(k_up_pair, d - 1, 64);
        if (d == 0) NBG_UP(k_up_internal, 0, 64);
    }
#undef NBG_UP
}

// Offsets of the sorted particles from their leaf's FP64 centre, rounded to FP32, with G*m:
// the grouped walk of an FP64 state differences these (plus FP32 offsets between leaf
// centres), so a pair difference never rounds an absolute position.
__global__ __launch_bounds__(256) void k_rel(const int32_t* __restrict__ leaf_of, const uint32_t* __restrict__ perm,
                                             const double* __restrict__ nd_geom, const float4* __restrict__ spos,
                                             const double* __restrict__ x64, const double* __restrict__ y64,
                                             const double* __restrict__ z64, int n, float4* __restrict__ srel) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int lf = leaf_of[p];
    const uint32_t i = perm[p];
    srel[p] = make_float4(static_cast<float>(x64[i] - nd_geom[4 * lf + 0]), static_cast<float>(y64[i] - nd_geom[4 * lf + 1]),
                          static_cast<float>(z64[i] - nd_geom[4 * lf + 2]), spos[p].w);
}

void launch_rel(nbg_ctx* c) {
    hipLaunchKernelGGL(k_rel, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->leaf_of, c->perm, c->nd_geom,
                       c->spos, c->x64, c->y64, c->z64, c->n, c->srel);
}

void launch_pack4(nbg_ctx* c) {
    // (an FP32 state's records are written by the key kernel)
    if (!c->f64_state) return;
    hipLaunchKernelGGL(k_pack4, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->x, c->y, c->z, c->m, c->n,
                       c->stage4);
}

}  // namespace nbg
