// nbg_common.hpp -- device helpers shared by the kernels (wave64 primitives, key math).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nbg {

__device__ __forceinline__ int lane_id() {
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
}

// Number of set bits of `m` in lanes below this one.
__device__ __forceinline__ int mbcnt(uint64_t m) {
    return static_cast<int>(
        __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Lane `l`'s value of v, for a wave-uniform l: one v_readlane into a scalar register (a
// lane shuffle goes through the LDS crossbar and waits on its latency).  Reads the lane's
// register whatever the execution mask.
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint32_t rdlane(uint32_t v, int l) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
}
__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint64_t rdlane(uint64_t v, int l) {
    return (static_cast<uint64_t>(rdlane(static_cast<uint32_t>(v >> 32), l)) << 32) | rdlane(static_cast<uint32_t>(v), l);
}
// v_writelane: v with lane l's value replaced by the wave-uniform w (l wave-uniform).  The
// LLVM intrinsic has no clang builtin here; it is reached through its symbol name.
extern "C" __device__ int nbg_llvm_writelane(int w, int l, int v) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ int wrlane(int v, int w, int l) { return nbg_llvm_writelane(w, l, v); }
__device__ __forceinline__ uint64_t wrlane(uint64_t v, uint64_t w, int l) {
    const uint32_t lo = static_cast<uint32_t>(wrlane(static_cast<int>(static_cast<uint32_t>(v)), static_cast<int>(static_cast<uint32_t>(w)), l));
    const uint32_t hi = static_cast<uint32_t>(
        wrlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), static_cast<int>(static_cast<uint32_t>(w >> 32)), l));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Orders LDS traffic of one wavefront: every LDS write issued before is visible to
// LDS reads issued after, for all lanes of the wave.  Waves of a workgroup never
// share LDS regions in the kernels that use this.
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); __builtin_amdgcn_wave_barrier(); }

// Row (16-lane) reductions through DPP: each step is one VALU op with a lane-permuted
// operand (quad xor 1, quad xor 2, half-row mirror, row mirror), no LDS round trip; every lane
// of a row ends with the row's result, and the wave's is the combination of lanes 15, 31,
// 47 and 63 read as scalars.
__device__ __forceinline__ uint32_t dpp_or_row(uint32_t v) {
    v |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false));
    v |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xF, 0xF, false));
    v |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xF, 0xF, false));
    v |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x140, 0xF, 0xF, false));
    return v;
}
__device__ __forceinline__ int dpp_add_row(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
    return v;
}
// OR of v over the wave (wave-uniform result)
__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
    const uint32_t lo = dpp_or_row(static_cast<uint32_t>(v)), hi = dpp_or_row(static_cast<uint32_t>(v >> 32));
    const uint32_t rlo = rdlane(lo, 15) | rdlane(lo, 31) | rdlane(lo, 47) | rdlane(lo, 63);
    const uint32_t rhi = rdlane(hi, 15) | rdlane(hi, 31) | rdlane(hi, 47) | rdlane(hi, 63);
    return (static_cast<uint64_t>(rhi) << 32) | rlo;
}
// sum of v over the wave (wave-uniform result)
__device__ __forceinline__ int wave_add(int v) {
    v = dpp_add_row(v);
    return rdlane(v, 15) + rdlane(v, 31) + rdlane(v, 47) + rdlane(v, 63);
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Order-preserving float <-> uint map (for atomic min/max of floats).
__device__ __forceinline__ uint32_t f2ord(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
}

// Common-prefix length, in levels, of two keys.  A short key holds levels 1..21 in
// bits 62..0 (3 bits per level, level 1 highest); a deep key adds levels 22..40 in
// bits 56..0 of the low word.  Equal keys return the number of levels held.
__device__ __forceinline__ int cpl_short(uint64_t a, uint64_t b) {
    const uint64_t d = a ^ b;
    if (d == 0) return 21;
    return (__clzll(static_cast<long long>(d)) - 1) / 3;
}
__device__ __forceinline__ int cpl_deep(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
    const uint64_t d = ah ^ bh;
    if (d != 0) return (__clzll(static_cast<long long>(d)) - 1) / 3;
    const uint64_t e = al ^ bl;
    if (e == 0) return 40;
    return 21 + (__clzll(static_cast<long long>(e)) - 7) / 3;
}

// Octant digit of level `d` (1-based) of a key.
__device__ __forceinline__ int key_digit(uint64_t hi, uint64_t lo, int d) {
    if (d <= 21) return static_cast<int>((hi >> (3 * (21 - d))) & 7u);
    return static_cast<int>((lo >> (3 * (40 - d))) & 7u);
}

}  // namespace nbg
