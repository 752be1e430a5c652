// nbg_pair.hpp -- pair kernels and load helpers shared by the force kernels (gfx950).
//
// Pair kernels (FP32):
//   Plummer  AccumPairPlummer   projects/nbody_core/src/Solvers.cpp:19-38
//   spline   AccumPairSpline    projects/nbody_core/src/Solvers.cpp:43-63 with
//            SplineChi3D        projects/nbody_core/include/ngrav/Softening.hpp:39-45
//   monopole M2P = the same pair kernel with the cell's G*mass at its centre of mass
//   quadrupole                  projects/nbody_core/src/Solvers.cpp:147-179 (unsoftened)
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/nbg.h"

namespace nbg {

// Wave-uniform reads of data no kernel writes while a force kernel runs (tree and sorted
// particles): going through the constant address space lets the compiler issue them as
// scalar (SMEM) loads into SGPRs, one per wave instead of one per lane.
typedef __attribute__((address_space(4))) const float cfloat;
typedef __attribute__((address_space(4))) const int32_t cint32;
typedef __attribute__((address_space(4))) const uint32_t cuint32;
__device__ __forceinline__ float4 cld(const float4* p, int i) {
    cfloat* q = (cfloat*)(p + i);
    return make_float4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ int32_t cld(const int32_t* p, int i) { return ((cint32*)p)[i]; }
typedef __attribute__((address_space(4))) const double cdouble;
__device__ __forceinline__ double4 cld(const double4* p, int i) {
    cdouble* q = (cdouble*)(p + i);
    return make_double4(q[0], q[1], q[2], q[3]);
}

// Records of a float4 array through a raw buffer resource: a wave-uniform index goes into
// the load's SGPR offset (index << 4 in 32 bits: valid for index < 2^28, which nbg_create's
// particle cap guarantees for sorted particles) and the record lands in VGPRs, so a group
// of such loads is in flight together without holding four SGPRs per record (scalar loads
// of a P2P group were split into three waits by SGPR pressure).  The third descriptor word
// is the gfx9-family default (32-bit data format).
typedef float f32x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rec_rsrc(const float4* p, int n_records) {
    const uint32_t bytes = static_cast<uint32_t>(n_records) * 16u;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(p), 0, static_cast<int>(bytes), 0x00020000);
}
__device__ __forceinline__ float4 bld(__amdgpu_buffer_rsrc_t r, int i) {
    const f32x4v t = __builtin_amdgcn_raw_buffer_load_b128(r, 0, static_cast<int>(static_cast<uint32_t>(i) << 4), 0);
    return make_float4(t[0], t[1], t[2], t[3]);
}
// The same for a per-lane index (a VGPR offset): each lane loads its own record.
__device__ __forceinline__ float4 bld_lane(__amdgpu_buffer_rsrc_t r, int i) {
    const f32x4v t = __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(static_cast<uint32_t>(i) << 4), 0, 0);
    return make_float4(t[0], t[1], t[2], t[3]);
}
__device__ __forceinline__ uint32_t cld_u32(const void* p, int i) { return ((cuint32*)p)[i]; }

template <bool KAHAN>
__device__ __forceinline__ void kadd(float& s, float& c, float v) {
    if (KAHAN) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    } else {
        s += v;
    }
}

// Softening constants of one evaluation.  Spline: h = sqrt(eps2) (Softening::Spline(h)
// stores eps = h, eps2 = h^2; sqrt of a correctly rounded square returns h exactly).
struct Soft {
    float eps2;    // Plummer eps^2
    float inv_h;   // spline 1/h
    float inv_h3;  // spline 1/h^3
    float four_h2; // spline (2h)^2: the kernel's support
};

inline Soft make_soft(const nbg_params& p) {
    Soft s;
    s.eps2 = static_cast<float>(p.eps2);
    const double h = std::sqrt(p.eps2);
    const double ih = (h > 0.0) ? 1.0 / h : 0.0;
    s.inv_h = static_cast<float>(ih);
    s.inv_h3 = static_cast<float>(ih * ih * ih);
    s.four_h2 = static_cast<float>(4.0 * p.eps2);
    return s;
}

// Coefficient f of a pair: the acceleration of a target from a source of G*m = gm at
// displacement d = source - target (r2 = d.d, unsoftened) is f * d.
//   Plummer: f = gm / (r^2 + eps^2)^(3/2).
//   Spline:  f = gm chi(q) / r^3, q = r/h, chi the enclosed-mass fraction of the cubic
//            spline kernel: chi = 1 for q >= 2; 1 - t^4 + 0.8 t^5 - t^6/6 (t = 2 - q) for
//            1 <= q < 2; q^3 (4/3 - 1.2 q^2 + 0.5 q^3) for q < 1, where chi / r^3 =
//            (4/3 - 1.2 q^2 + 0.5 q^3) / h^3 needs no division and stays finite at r = 0
//            (d = 0 there, so the term vanishes as the reference's r <= 1e-300 branch makes it).
//            Most pairs of a walk lie beyond 2h: the kernel branch runs only when a lane of
//            the wave is inside the support (a wave-uniform branch on a ballot).
template <bool SPLINE>
__device__ __forceinline__ float pair_coef(float r2, float gm, const Soft& s) {
    if (!SPLINE) {
        const float inv = __builtin_amdgcn_rsqf(r2 + s.eps2);
        return gm * (inv * inv * inv);
    } else {
        const float inv = __builtin_amdgcn_rsqf(r2);  // inf at r2 = 0, replaced below
        float f = gm * (inv * inv * inv);              // q >= 2: Newtonian
        const bool inside = r2 < s.four_h2;
        if (__ballot(inside)) {
            const float q = sqrtf(r2) * s.inv_h;
            const float t = 2.f - q;
            const float t2 = t * t;
            const float chi = 1.f - (t2 * t2) * (1.f - t * (0.8f - t * (1.f / 6.f)));
            const float f_in = gm * s.inv_h3 * (4.f / 3.f + (q * q) * (0.5f * q - 1.2f));
            f = inside ? ((q < 1.f) ? f_in : f * chi) : f;
        }
        return f;
    }
}

// pair_coef of the displacement (dx, dy, dz); for Plummer the softening enters the
// fused multiply-add chain of r^2 + eps^2 (one instruction less than adding it after).
template <bool SPLINE>
__device__ __forceinline__ float pair_coef_d(float dx, float dy, float dz, float gm, const Soft& s) {
    if (!SPLINE) {
        const float inv = __builtin_amdgcn_rsqf(fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, s.eps2))));
        return gm * (inv * inv * inv);
    } else {
        return pair_coef<true>(dx * dx + dy * dy + dz * dz, gm, s);
    }
}

// M2P of one cell (com + G*mass in cm, G*(Qxx,Qyy,Qxy,Qxz) in q, G*Qyz in qyz) onto a
// target at (tx,ty,tz): softened monopole plus the unsoftened traceless quadrupole.
template <bool QUAD, bool SPLINE>
__device__ __forceinline__ void m2p_cell(float4 cm, float4 q, float qyz, float tx, float ty, float tz,
                                         const Soft& soft, float& ex, float& ey, float& ez) {
    const float dx = cm.x - tx, dy = cm.y - ty, dz = cm.z - tz;
    const float dist2 = dx * dx + dy * dy + dz * dz;
    const float f = pair_coef<SPLINE>(dist2, cm.w, soft);
    ex = f * dx;
    ey = f * dy;
    ez = f * dz;
    if (QUAD) {
        const float qxx = q.x, qyy = q.y, qxy = q.z, qxz = q.w, qzz = -(qxx + qyy);
        const float qdx = qxx * dx + qxy * dy + qxz * dz;
        const float qdy = qxy * dx + qyy * dy + qyz * dz;
        const float qdz = qxz * dx + qyz * dy + qzz * dz;
        const float qd2 = dx * qdx + dy * qdy + dz * qdz;
        const float ir = __builtin_amdgcn_rsqf(dist2);
        const float ir2 = ir * ir;
        const float ir5 = ir2 * ir2 * ir;
        const float c2 = 2.5f * qd2 * (ir5 * ir2);
        ex += c2 * dx - ir5 * qdx;
        ey += c2 * dy - ir5 * qdy;
        ez += c2 * dz - ir5 * qdz;
    }
}

// m2p_cell's field of one cell added into (sx,sy,sz) when `use`, nothing otherwise.  The
// terms are grouped per component as (f + c2) d - ir5 Qd, so a lane spends two fused
// multiply-adds per component.  A lane that does not use the cell adds +inf to its r^2:
// both reciprocal roots are then +0 and every term an exact zero, also where its position
// is the cell's com (r^2 = 0), so no select is needed (one instruction less per cell than
// two selects of the results).
// q2 = (Qyz, Qzz), Qzz = -(Qxx + Qyy) stored with the cell's record.
// r2add: the addend of r^2, 0 (the cell is used) or +inf (not used).
template <bool QUAD, bool SPLINE>
__device__ __forceinline__ void m2p_add_r2(float4 cm, float4 q, float2 q2, float tx, float ty, float tz,
                                           const Soft& soft, float r2add, float& sx, float& sy, float& sz) {
    const float dx = cm.x - tx, dy = cm.y - ty, dz = cm.z - tz;
    const float dist2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, r2add)));
    float g = pair_coef<SPLINE>(dist2, cm.w, soft);
    if (QUAD) {
        const float qxx = q.x, qyy = q.y, qxy = q.z, qxz = q.w, qyz = q2.x, qzz = q2.y;
        const float qdx = qxx * dx + qxy * dy + qxz * dz;
        const float qdy = qxy * dx + qyy * dy + qyz * dz;
        const float qdz = qxz * dx + qyz * dy + qzz * dz;
        const float qd2 = dx * qdx + dy * qdy + dz * qdz;
        const float ir = __builtin_amdgcn_rsqf(dist2);
        const float ir2 = ir * ir;
        const float ir5 = ir2 * ir2 * ir;
        g = fmaf(ir5, (2.5f * qd2) * ir2, g);
        sx = fmaf(g, dx, sx);
        sy = fmaf(g, dy, sy);
        sz = fmaf(g, dz, sz);
        sx = fmaf(-ir5, qdx, sx);
        sy = fmaf(-ir5, qdy, sy);
        sz = fmaf(-ir5, qdz, sz);
    } else {
        sx = fmaf(g, dx, sx);
        sy = fmaf(g, dy, sy);
        sz = fmaf(g, dz, sz);
    }
}
template <bool QUAD, bool SPLINE>
__device__ __forceinline__ void m2p_add(float4 cm, float4 q, float2 q2, float tx, float ty, float tz,
                                        const Soft& soft, bool use, float& sx, float& sy, float& sz) {
    m2p_add_r2<QUAD, SPLINE>(cm, q, q2, tx, ty, tz, soft, use ? 0.f : __builtin_inff(), sx, sy, sz);
}

// `yes` where bit `sh` of the mask word w is set, `no` where it is clear: a signed one-bit
// field extract (0 or all ones) and a bitfield select, two instructions.
__device__ __forceinline__ float mask_pick(uint32_t w, int sh, float yes, float no) {
    const uint32_t b = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(w), sh, 1));
    return __uint_as_float((b & __float_as_uint(yes)) | (~b & __float_as_uint(no)));
}

}  // namespace nbg
