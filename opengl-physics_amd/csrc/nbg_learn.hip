// nbg_learn.hip -- learned far-field correction: moments, features, BF16 MFMA inference
// and training kernels, sample generation from direct-sum snapshots.
//
// The model is a perceptron 13 -> 64 -> 64 -> 64 -> 3 with SiLU.  One evaluation unit is
// an accepted (leaf group, cell) pair listed by the walk.  Its features describe the cell
// as seen from the group's centre of mass, in units of the cell:
//   u (3)        unit vector from the cell's com to the group's com
//   s/r          cell size over distance
//   a3 (3)       octupole force shape: 7.5 S.uu - 17.5 (S:uuu) u - 1.5 t + 7.5 (t.u) u,
//                S = third moments / (M s^3), t_i = S_ijj
//   Q.u (3)      second moments / (M s^2), contracted with u
//   R_g/r        group radius over distance
//   eps/r        softening over distance
//   tr Q         trace of the second moments / (M s^2)
// and its output, times the scale G M s^3 / r^5, is the residual of the cell's monopole +
// quadrupole force against direct summation, averaged over the group's members.
//
// Matrix-core layout (v_mfma_f32_16x16x16_bf16, 64 lanes, g = lane / 16, c = lane % 16):
//   A operand  lane holds A[row c][k 4g .. 4g+3]
//   B operand  lane holds B[k 4g .. 4g+3][col c]
//   C/D        lane holds D[row 4g .. 4g+3][col c]
// Layers are computed transposed (hidden x pairs), so the D registers of one layer are
// the B operand of the next with no data movement: D rows 4g+j of row block mb are the
// B rows 4g+j of K block mb.
#include <cstring>
#include <type_traits>
#include <vector>

#include "nbg_common.hpp"
#include "nbg_internal.hpp"

namespace nbg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 hbf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kF = NBG_LEARN_FEATURES;  // 13
constexpr int kH = NBG_LEARN_HIDDEN;    // 64
// trainable blob offsets (after the 32 standardisation entries)
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kF * kH;   // 832
constexpr int kOffW2 = kOffB1 + kH;        // 896
constexpr int kOffB2 = kOffW2 + kH * kH;   // 4992
constexpr int kOffW3 = kOffB2 + kH;        // 5056
constexpr int kOffB3 = kOffW3 + kH * kH;   // 9152
constexpr int kOffW4 = kOffB3 + kH;        // 9216
constexpr int kOffB4 = kOffW4 + kH * 3;    // 9408
constexpr int kNTrain = kOffB4 + 3;        // 9411
static_assert(kNTrain == NBG_LEARN_NTRAIN, "blob layout");
static_assert(NBG_LEARN_ACT_SLOT >= NBG_LEARN_FEATURES && NBG_LEARN_ACT_SLOT < 16, "activation slot is an unused mean");
constexpr int kNorm = 32;                  // mean[16], 1/std[16]
constexpr int kActSilu = NBG_LEARN_ACT_SILU;
constexpr int kActRelu = NBG_LEARN_ACT_RELU;

__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t b = __float_as_uint(f);
    b += 0x7fffu + ((b >> 16) & 1u);
    return static_cast<uint16_t>(b >> 16);
}
__device__ __forceinline__ float bf16_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
// two floats -> two bf16 (round to nearest even) in one v_cvt_pk_bf16_f32; low half = a
__device__ __forceinline__ uint32_t bf16x2_rne(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, hbf16x2));
}

__device__ __forceinline__ float sigmoid_f(float z) {
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }
// The network's activation, chosen by the model (kActSilu, kActRelu), of the BF16-rounded
// pre-activation, rounded to BF16.  SiLU inference reads it from a 64K-entry table indexed
// by those BF16 bits; ReLU is exact on the BF16 bits (a negative value, -0 included, has
// its sign bit set and becomes +0).
__device__ __forceinline__ uint16_t act_bits(uint16_t zb) { return bf16_rne(silu_f(bf16_f32(zb))); }
template <int ACT>
__device__ __forceinline__ uint16_t act_of(float z) {
    const uint16_t b = bf16_rne(z);
    if (ACT == kActRelu) return (b & 0x8000u) ? uint16_t{0} : b;
    return act_bits(b);
}
__device__ __forceinline__ float dsilu_f(float z) {
    const float s = sigmoid_f(z);
    return s * (1.f + z * (1.f - s));
}
// The activation's derivative at the unrounded pre-activation (training).
template <int ACT>
__device__ __forceinline__ float dact(float z) {
    if (ACT == kActRelu) return z > 0.f ? 1.f : 0.f;
    return dsilu_f(z);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// K = 32: lane holds A[row c][k 8g .. 8g+7], B[k 8g .. 8g+7][col c]; D as for K = 16.
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 cat8(bf16x4 lo, bf16x4 hi) {
    const uint64_t l = __builtin_bit_cast(uint64_t, lo), h = __builtin_bit_cast(uint64_t, hi);
    struct P {
        uint64_t a, b;
    } v{l, h};
    return __builtin_bit_cast(bf16x8, v);
}

// ---------------------------------------------------------------------------- features
// tgt: group com (xyz) and radius (w); cm: cell com (xyz) and G*M (w); size2 = (2h)^2;
// mom: the cell's four moment words.  Writes the 13 raw features and the output scale.
__device__ __forceinline__ void pair_features(float4 tgt, float4 cm, float size2, const float4* mom, float eps,
                                              float f[kF], float& scale) {
    const float Rx = tgt.x - cm.x, Ry = tgt.y - cm.y, Rz = tgt.z - cm.z;
    const float r2 = Rx * Rx + Ry * Ry + Rz * Rz;
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ux = Rx * ir, uy = Ry * ir, uz = Rz * ir;
    const float s = sqrtf(size2);
    const float M = cm.w;
    const float i2 = 1.f / (M * size2);
    const float i3 = i2 / s;
    const float4 m0 = mom[0], m1 = mom[1], m2 = mom[2], m3 = mom[3];
    const float sxx = m0.x, syy = m0.y, szz = m0.z, sxy = m0.w, sxz = m1.x, syz = m1.y;
    const float xxx = m1.z, yyy = m1.w, zzz = m2.x, xxy = m2.y, xxz = m2.z, xyy = m2.w, yyz = m3.x, xzz = m3.y,
                yzz = m3.z, xyz = m3.w;
    const float uxx = ux * ux, uyy = uy * uy, uzz = uz * uz, uxy = 2.f * ux * uy, uxz = 2.f * ux * uz,
                uyz = 2.f * uy * uz;
    const float Sx = (xxx * uxx + xyy * uyy + xzz * uzz + xxy * uxy + xxz * uxz + xyz * uyz) * i3;
    const float Sy = (xxy * uxx + yyy * uyy + yzz * uzz + xyy * uxy + xyz * uxz + yyz * uyz) * i3;
    const float Sz = (xxz * uxx + yyz * uyy + zzz * uzz + xyz * uxy + xzz * uxz + yzz * uyz) * i3;
    const float tx = (xxx + xyy + xzz) * i3, ty = (xxy + yyy + yzz) * i3, tz = (xxz + yyz + zzz) * i3;
    const float Suuu = Sx * ux + Sy * uy + Sz * uz;
    const float tu = tx * ux + ty * uy + tz * uz;
    const float cu = 7.5f * tu - 17.5f * Suuu;
    f[0] = ux;
    f[1] = uy;
    f[2] = uz;
    f[3] = s * ir;
    f[4] = 7.5f * Sx - 1.5f * tx + cu * ux;
    f[5] = 7.5f * Sy - 1.5f * ty + cu * uy;
    f[6] = 7.5f * Sz - 1.5f * tz + cu * uz;
    f[7] = (sxx * ux + sxy * uy + sxz * uz) * i2;
    f[8] = (sxy * ux + syy * uy + syz * uz) * i2;
    f[9] = (sxz * ux + syz * uy + szz * uz) * i2;
    f[10] = tgt.w * ir;
    f[11] = eps * ir;
    f[12] = (sxx + syy + szz) * i2;
    const float ir2 = ir * ir;
    scale = M * (s * size2) * (ir2 * ir2 * ir);
}

// ---------------------------------------------------------------------------- weights
// The per-lane weight fragments of the whole network (A operands) and its biases (as
// C initial values), held in registers for the life of a wave.
struct Frags {
    bf16x4 a1[4], a2[4][4], a3[4][4], a4[4];
    f32x4 b1[4], b2[4], b3[4], b4;
};

__device__ __forceinline__ bf16x4 ld4_bf16(const uint16_t* w, int i0, int stride, bool valid) {
    bf16x4 v;
    v[0] = valid ? static_cast<short>(w[i0]) : 0;
    v[1] = valid ? static_cast<short>(w[i0 + stride]) : 0;
    v[2] = valid ? static_cast<short>(w[i0 + 2 * stride]) : 0;
    v[3] = valid ? static_cast<short>(w[i0 + 3 * stride]) : 0;
    return v;
}

__device__ __forceinline__ void load_frags(Frags& F, const uint16_t* __restrict__ w16, const float* __restrict__ b32,
                                           int lane) {
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        // A1[h][f] = W1[f][h]: h = 16 mb + c, f = 4g + j (< 13)
        bf16x4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 4 * g + j;
            v[j] = (f < kF) ? static_cast<short>(w16[kOffW1 + f * kH + 16 * mb + c]) : 0;
        }
        F.a1[mb] = v;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            // A[h][i] = W[i][h]: h = 16 mb + c, i = 16 kb + 4g + j
            F.a2[mb][kb] = ld4_bf16(w16, kOffW2 + (16 * kb + 4 * g) * kH + 16 * mb + c, kH, true);
            F.a3[mb][kb] = ld4_bf16(w16, kOffW3 + (16 * kb + 4 * g) * kH + 16 * mb + c, kH, true);
        }
        // A4[o][i] = W4[i][o]: o = c (< 3), i = 16 mb + 4g + j
        F.a4[mb] = ld4_bf16(w16, kOffW4 + (16 * mb + 4 * g) * 3 + (c < 3 ? c : 0), 3, c < 3);
        f32x4 b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = b32[kOffB1 + 16 * mb + 4 * g + j];
        F.b1[mb] = b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = b32[kOffB2 + 16 * mb + 4 * g + j];
        F.b2[mb] = b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = b32[kOffB3 + 16 * mb + 4 * g + j];
        F.b3[mb] = b;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) F.b4[j] = (4 * g + j < 3) ? b32[kOffB4 + 4 * g + j] : 0.f;
}

// Activation of a D fragment, packed as the next layer's B fragment.
template <int ACT>
struct ActDirect {  // computed (training)
    __device__ __forceinline__ bf16x4 operator()(f32x4 z) const {
        bf16x4 h;
#pragma unroll
        for (int j = 0; j < 4; ++j) h[j] = static_cast<short>(act_of<ACT>(z[j]));
        return h;
    }
};
// Inference with ReLU: one v_cvt_pk_bf16_f32 and one v_pk_max_i16 per two values (the
// BF16 bits of a negative value, -0 included, are a negative 16-bit integer).
struct ActRelu {
    __device__ __forceinline__ bf16x4 operator()(f32x4 z, int) const {
        typedef short s16x2 __attribute__((ext_vector_type(2)));
        const s16x2 zero = {0, 0};
        const s16x2 lo = __builtin_elementwise_max(__builtin_bit_cast(s16x2, bf16x2_rne(z[0], z[1])), zero);
        const s16x2 hi = __builtin_elementwise_max(__builtin_bit_cast(s16x2, bf16x2_rne(z[2], z[3])), zero);
        return __builtin_bit_cast(bf16x4, (static_cast<uint64_t>(__builtin_bit_cast(uint32_t, hi)) << 32) |
                                              __builtin_bit_cast(uint32_t, lo));
    }
};
// Inference: the same function, three of the four row blocks looked up in a 64K-entry LDS
// table and one computed, so the LDS pipe (bank-conflicted lookups) and the transcendental
// VALU work proceed side by side (measured best of 1/4, 1/2, 3/4, 4/4 looked up); both
// give identical bits.
struct ActTable {
    const uint16_t* tab;
    __device__ __forceinline__ bf16x4 lookup(f32x4 z) const {
        const uint32_t lo = bf16x2_rne(z[0], z[1]);
        const uint32_t hi = bf16x2_rne(z[2], z[3]);
        const uint32_t a = tab[lo & 0xffffu], b = tab[lo >> 16], c = tab[hi & 0xffffu], d = tab[hi >> 16];
        return __builtin_bit_cast(bf16x4, (static_cast<uint64_t>(d << 16 | c) << 32) | (b << 16 | a));
    }
    __device__ __forceinline__ bf16x4 compute(f32x4 z) const {
        const uint32_t lo = bf16x2_rne(z[0], z[1]);
        const uint32_t hi = bf16x2_rne(z[2], z[3]);
        const float s0 = silu_f(__uint_as_float(lo << 16)), s1 = silu_f(__uint_as_float(lo & 0xffff0000u));
        const float s2 = silu_f(__uint_as_float(hi << 16)), s3 = silu_f(__uint_as_float(hi & 0xffff0000u));
        const uint32_t r0 = bf16x2_rne(s0, s1), r1 = bf16x2_rne(s2, s3);
        return __builtin_bit_cast(bf16x4, (static_cast<uint64_t>(r1) << 32) | r0);
    }
    __device__ __forceinline__ bf16x4 operator()(f32x4 z, int mb) const { return (mb == 3) ? compute(z) : lookup(z); }
};

// One 16-pair tile through the network.  x: this lane's B fragment of the standardised
// features (rows 4g..4g+3 of column c).  Returns D of the output layer (lanes 0..15 hold
// outputs 0..2 of pair c).  With Z != nullptr the pre-activations of the three hidden
// layers are stored as Z[layer][pair][hidden] (pair = col0 + c).
template <bool STORE, class Act>
__device__ __forceinline__ f32x4 mlp_tile(const Frags& F, bf16x4 x, float* Z, int col0, int lane, const Act& act) {
    const int g = lane >> 4, c = lane & 15;
    bf16x4 h[4];
    f32x4 z[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) z[mb] = mfma16(F.a1[mb], x, F.b1[mb]);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        if (STORE) *reinterpret_cast<f32x4*>(Z + (col0 + c) * kH + 16 * mb + 4 * g) = z[mb];
        h[mb] = act(z[mb]);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        f32x4 acc = F.b2[mb];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma16(F.a2[mb][kb], h[kb], acc);
        z[mb] = acc;
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        if (STORE) *reinterpret_cast<f32x4*>(Z + 64 * kH + (col0 + c) * kH + 16 * mb + 4 * g) = z[mb];
        h[mb] = act(z[mb]);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        f32x4 acc = F.b3[mb];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma16(F.a3[mb][kb], h[kb], acc);
        z[mb] = acc;
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        if (STORE) *reinterpret_cast<f32x4*>(Z + 2 * 64 * kH + (col0 + c) * kH + 16 * mb + 4 * g) = z[mb];
        h[mb] = act(z[mb]);
    }
    f32x4 y = F.b4;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) y = mfma16(F.a4[kb], h[kb], y);
    return y;
}

// ---------------------------------------------------------------------------- inference
// Inference fragments for the K = 32 instruction.  The hidden layers' D registers (rows
// 4g..4g+3 of row blocks 2kb and 2kb+1) are concatenated into the B operand of K block
// kb, so K position 8g + j of block kb is hidden unit 16 (2 kb + j / 4) + 4 g + j % 4;
// the weight fragments below are gathered in that order.  The first layer (K = 13 -> 16)
// keeps the K = 16 instruction.
struct Frags32 {
    bf16x4 a1[4];
    bf16x8 a2[4][2], a3[4][2], a4[2];
    f32x4 b1[4], b2[4], b3[4], b4;
};

__device__ __forceinline__ int hid32(int kb, int g, int j) { return 16 * (2 * kb + (j >> 2)) + 4 * g + (j & 3); }

__device__ __forceinline__ void load_frags32(Frags32& F, const uint16_t* __restrict__ w16,
                                             const float* __restrict__ b32, int lane) {
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        bf16x4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 4 * g + j;
            v[j] = (f < kF) ? static_cast<short>(w16[kOffW1 + f * kH + 16 * mb + c]) : 0;
        }
        F.a1[mb] = v;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            uint16_t u2[8], u3[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int i = hid32(kb, g, j);
                u2[j] = w16[kOffW2 + i * kH + 16 * mb + c];
                u3[j] = w16[kOffW3 + i * kH + 16 * mb + c];
            }
            F.a2[mb][kb] = __builtin_bit_cast(bf16x8, u2);
            F.a3[mb][kb] = __builtin_bit_cast(bf16x8, u3);
        }
        f32x4 b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = b32[kOffB1 + 16 * mb + 4 * g + j];
        F.b1[mb] = b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = b32[kOffB2 + 16 * mb + 4 * g + j];
        F.b2[mb] = b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = b32[kOffB3 + 16 * mb + 4 * g + j];
        F.b3[mb] = b;
    }
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        uint16_t u4[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) u4[j] = (c < 3) ? w16[kOffW4 + hid32(kb, g, j) * 3 + c] : 0;
        F.a4[kb] = __builtin_bit_cast(bf16x8, u4);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) F.b4[j] = (4 * g + j < 3) ? b32[kOffB4 + 4 * g + j] : 0.f;
}

template <class Act>
__device__ __forceinline__ void mlp32_tile2(const Frags32& F, bf16x4 xa, bf16x4 xb, const Act& act, f32x4& ya,
                                            f32x4& yb) {
    bf16x4 ha[4], hb[4];
    f32x4 za[4], zb[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        za[mb] = mfma16(F.a1[mb], xa, F.b1[mb]);
        zb[mb] = mfma16(F.a1[mb], xb, F.b1[mb]);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        ha[mb] = act(za[mb], mb);
        hb[mb] = act(zb[mb], mb);
    }
#pragma unroll
    for (int L = 0; L < 2; ++L) {
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const f32x4 bias = L == 0 ? F.b2[mb] : F.b3[mb];
            f32x4 a = bias, b = bias;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const bf16x8 w = L == 0 ? F.a2[mb][kb] : F.a3[mb][kb];
                a = mfma32(w, cat8(ha[2 * kb], ha[2 * kb + 1]), a);
                b = mfma32(w, cat8(hb[2 * kb], hb[2 * kb + 1]), b);
            }
            za[mb] = a;
            zb[mb] = b;
        }
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            ha[mb] = act(za[mb], mb);
            hb[mb] = act(zb[mb], mb);
        }
    }
    f32x4 a = F.b4, b = F.b4;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        a = mfma32(F.a4[kb], cat8(ha[2 * kb], ha[2 * kb + 1]), a);
        b = mfma32(F.a4[kb], cat8(hb[2 * kb], hb[2 * kb + 1]), b);
    }
    ya = a;
    yb = b;
}

// ---------------------------------------------------------------------------- fixed point
// v * 2^e rounded to the nearest integer (ties to even) as a two's-complement 64-bit word:
// the product is exact in FP32 (|v 2^e| < 2^49 is checked; an underflow rounds to 0 as the
// exact value would), and r - 2^32 floor(r / 2^32) keeps at most the 24 significant bits of
// r below bit 32, so it is exact too.  Equal to rounding the FP64 product.
__device__ __forceinline__ uint64_t fix_term(float v, int e, bool& over) {
    const float t = ldexpf(v, e);
    over |= !(fabsf(t) < 0x1p49f);
    const float r = __builtin_rintf(t);
    const float hi = floorf(r * 0x1p-32f);
    const float lo = fmaf(hi, -0x1p32f, r);
    return (static_cast<uint64_t>(static_cast<uint32_t>(static_cast<int32_t>(hi))) << 32) |
           static_cast<uint64_t>(static_cast<uint32_t>(lo));
}

// DPP move with control CTRL over the rows in ROWS: a lane whose source lies outside the
// wave (or whose row is not in ROWS) keeps `fill`.  Row shifts right (0x111..0x11f) move
// within rows of 16 lanes; row_bcast15 (0x142, rows 1 and 3) and row_bcast31 (0x143, rows
// 2 and 3) bring the last lane of the row(s) before.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v, uint32_t fill) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(fill), static_cast<int>(v), CTRL, ROWS, 0xf, false));
}
// One step of an inclusive segmented sum over the wave (Hillis-Steele on DPP moves): a lane
// adds the partial sums of its DPP source lane when both lie in the same run.
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_sum_step(int seg, uint64_t& a, uint64_t& b, uint64_t& c) {
    const bool same = static_cast<int>(dpp_mov<CTRL, ROWS>(static_cast<uint32_t>(seg), 0xffffffffu)) == seg;
    auto mv = [](uint64_t v) {
        return (static_cast<uint64_t>(dpp_mov<CTRL, ROWS>(static_cast<uint32_t>(v >> 32), 0u)) << 32) |
               dpp_mov<CTRL, ROWS>(static_cast<uint32_t>(v), 0u);
    };
    const uint64_t sa = mv(a), sb = mv(b), sc = mv(c);
    a += same ? sa : 0ull;
    b += same ? sb : 0ull;
    c += same ? sc : 0ull;
}
// Inclusive sums of (a, b, c) over each run of equal `seg` (runs are contiguous lanes):
// shifts by 1, 2, 4, 8 within rows, then the row broadcasts carry runs across rows.
__device__ __forceinline__ void wave_seg_sum(int seg, uint64_t& a, uint64_t& b, uint64_t& c) {
    seg_sum_step<0x111, 0xf>(seg, a, b, c);
    seg_sum_step<0x112, 0xf>(seg, a, b, c);
    seg_sum_step<0x114, 0xf>(seg, a, b, c);
    seg_sum_step<0x118, 0xf>(seg, a, b, c);
    seg_sum_step<0x142, 0xa>(seg, a, b, c);
    seg_sum_step<0x143, 0xc>(seg, a, b, c);
}

// ---------------------------------------------------------------------------- inference
// One persistent kernel per evaluation: a lane computes the standardised BF16 feature row
// of one listed pair (16 x 2 B, entries 13..15 zero) and its output scale from the node
// data, the wave stages its 64 rows in LDS and reads them back as matrix-core fragments.
struct InferArgs {
    const int2* pairs;
    const DevScalars* sc;
    DevScalars* scw;
    uint32_t cap;
    const float4* nd_rec;  // 4 per node: com, quad, (Qyz, Qzz, first particle), aux
    const float4* nd_mom;
    float eps;
    const uint16_t* w16;
    const float* w32;    // full blob (standardisation first)
    float4* pv;          // two-pass sums: per pair, the correction
    uint32_t* unit;      // two-pass sums: per leaf, largest |component| of its corrections (float bits)
    long long* acc;      // one-pass sums: per leaf, 4 words of fixed point at the evaluation's unit
};

constexpr int kInferWaves = 8;
constexpr int kInferCUs = 256;        // MI355X: 8 XCDs x 32 CUs

// The node words one pair's features are computed from.
struct PairRaw {
    float4 lc, cm, m0, m1, m2, m3;
    float lr, size2;
};

// Persistent: one 8-wave block per CU (the weight fragments held in registers take ~250
// VGPRs, so two waves per SIMD; with SiLU the 128 KB activation table fills most of the
// LDS as well).  A wave takes 64 pairs per step as four 16-pair tiles, two at a time.  The pair ids run two steps ahead and the node words one
// step ahead of the step computed.
template <bool TWO_PASS, int ACT>
__global__ __launch_bounds__(64 * kInferWaves) void k_learn_infer(InferArgs A) {
    __shared__ __attribute__((aligned(16))) uint16_t s_act[ACT == kActSilu ? 65536 : 2];
    __shared__ uint2 s_row[kInferWaves][64][4];
    if (ACT == kActSilu) {
        for (int i = threadIdx.x; i < 65536; i += blockDim.x) s_act[i] = act_bits(static_cast<uint16_t>(i));
        __syncthreads();
    }
    typedef typename std::conditional<ACT == kActSilu, ActTable, ActRelu>::type Act;
    Act act;
    if constexpr (ACT == kActSilu) act.tab = s_act;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = lane_id();
    const int g = lane >> 4, c = lane & 15;
    const uint32_t listed = A.sc->learn_pairs;
    const uint32_t np = min(listed, A.cap);
    if (blockIdx.x == 0 && threadIdx.x == 0 && listed > A.cap) atomicOr(&A.scw->sticky_learn, 1);
    const int64_t nchunks = (static_cast<int64_t>(np) + 63) / 64;
    const int64_t wid = static_cast<int64_t>(blockIdx.x) * kInferWaves + wave;
    const int64_t nw = static_cast<int64_t>(gridDim.x) * kInferWaves;
    if (wid >= nchunks) return;
    Frags32 F;
    load_frags32(F, A.w16, A.w32 + kNorm, lane);
    const float* mu = A.w32;
    const float* isd = A.w32 + 16;
    uint2(*row)[4] = s_row[wave];
    auto load_pair = [&](int64_t ch) {
        const int64_t q = ch * 64 + lane;
        return (ch < nchunks && q < np) ? A.pairs[q] : make_int2(-1, -1);
    };
    auto load_raw = [&](int2 pr, PairRaw& r) {
        if (pr.x >= 0) {
            r.lc = A.nd_rec[4 * pr.x];
            r.lr = A.nd_rec[4 * pr.x + 3].z;
            r.cm = A.nd_rec[4 * pr.y];
            r.size2 = A.nd_rec[4 * pr.y + 3].y;
            r.m0 = A.nd_mom[4 * pr.y + 0];
            r.m1 = A.nd_mom[4 * pr.y + 1];
            r.m2 = A.nd_mom[4 * pr.y + 2];
            r.m3 = A.nd_mom[4 * pr.y + 3];
        }
    };
    int2 pr_n2 = load_pair(wid + nw);
    int2 pr_n = load_pair(wid);
    PairRaw raw_n{};
    load_raw(pr_n, raw_n);
    for (int64_t ch = wid; ch < nchunks; ch += nw) {
        const int leaf = pr_n.x;
        const PairRaw raw = raw_n;
        pr_n = pr_n2;
        pr_n2 = load_pair(ch + 2 * nw);
        load_raw(pr_n, raw_n);
        // this lane's pair: features, standardised, as a BF16 row in LDS
        float scale = 0.f;
        {
            const float4 mom[4] = {raw.m0, raw.m1, raw.m2, raw.m3};
            float f[kF];
            pair_features(make_float4(raw.lc.x, raw.lc.y, raw.lc.z, raw.lr), raw.cm, raw.size2, mom, A.eps, f,
                          scale);
            float xv[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) xv[k] = (k < kF) ? (f[k] - mu[k]) * isd[k] : 0.f;
            uint32_t w[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) w[k] = leaf >= 0 ? bf16x2_rne(xv[2 * k], xv[2 * k + 1]) : 0u;
            scale = leaf >= 0 ? scale : 0.f;
#pragma unroll
            for (int k = 0; k < 4; ++k) row[lane][k] = make_uint2(w[2 * k], w[2 * k + 1]);
        }
        wave_sync();
        // fragment of tile t: pair 16 t + c, features 4g..4g+3
        uint2 x[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) x[t] = row[16 * t + c][g];
        wave_sync();
        float o0 = 0.f, o1 = 0.f, o2 = 0.f;
#pragma unroll
        for (int t = 0; t < 4; t += 2) {
            f32x4 ya, yb;
            mlp32_tile2(F, __builtin_bit_cast(bf16x4, x[t]), __builtin_bit_cast(bf16x4, x[t + 1]), act, ya, yb);
            // lanes 0..15 hold the outputs of pair 16 t + c; move them to lane 16 t + c
            const float a0 = __shfl(ya[0], c, 64), a1 = __shfl(ya[1], c, 64), a2 = __shfl(ya[2], c, 64);
            const float b0 = __shfl(yb[0], c, 64), b1 = __shfl(yb[1], c, 64), b2 = __shfl(yb[2], c, 64);
            if (g == t) {
                o0 = a0;
                o1 = a1;
                o2 = a2;
            }
            if (g == t + 1) {
                o0 = b0;
                o1 = b1;
                o2 = b2;
            }
        }
        const float v0 = scale * o0, v1 = scale * o1, v2 = scale * o2;
        if (TWO_PASS) {
            const int prev = __shfl_up(leaf, 1, 64);
            const bool head = lane == 0 || prev != leaf;
            const uint64_t heads = ballot(head);
            const int seg = __popcll(heads & ((2ull << lane) - 1ull));
            const int next_seg = __shfl_down(seg, 1, 64);
            const bool tail = lane == 63 || next_seg != seg;
            // each pair's correction to its slot; the largest |component| over each run of
            // equal leaf ids (the walk lists a leaf's pairs together) to the leaf, by one
            // atomic max per run: the maximum does not depend on how the runs are cut, so
            // the leaf's fixed-point unit is the same on every run
            const int64_t q = ch * 64 + lane;
            if (leaf >= 0) A.pv[q] = make_float4(v0, v1, v2, 0.f);
            float mx = fmaxf(fabsf(v0), fmaxf(fabsf(v1), fabsf(v2)));
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const float t = __shfl_up(mx, d, 64);
                const int ts = __shfl_up(seg, d, 64);
                if (lane >= d && ts == seg) mx = fmaxf(mx, t);
            }
            if (tail && leaf >= 0 && mx > 0.f) atomicMax(A.unit + leaf, __float_as_uint(mx));
        } else {
            // exact sums in 64-bit fixed point at the evaluation's unit (integer addition is
            // associative: the sums do not depend on the order in which the walk listed the
            // pairs), a segmented sum over each run of equal leaf ids (DPP moves: no LDS round
            // trips), one atomic triple per run; a term beyond the range flags the evaluation
            // for the two-pass sums
            const int e = A.sc->learn_fix_e;
            bool over = false;
            uint64_t i0 = fix_term(v0, e, over), i1 = fix_term(v1, e, over), i2 = fix_term(v2, e, over);
            if (over) {
                atomicOr(&A.scw->learn_fix_over, 1);
                atomicOr(&A.scw->sticky_learn, 1);
                i0 = i1 = i2 = 0;
            }
            // runs of equal leaf ids (lane 0 compares with the fill, which no leaf id equals)
            const int prev = __shfl_up(leaf, 1, 64);
            const uint64_t heads = ballot(lane == 0 || prev != leaf);
            const int seg = __popcll(heads & ((2ull << lane) - 1ull));
            const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
            wave_seg_sum(seg, i0, i1, i2);
            if (tail && leaf >= 0) {
                unsigned long long* dst = reinterpret_cast<unsigned long long*>(A.acc + 4 * static_cast<int64_t>(leaf));
                atomicAdd(dst + 0, static_cast<unsigned long long>(i0));
                atomicAdd(dst + 1, static_cast<unsigned long long>(i1));
                atomicAdd(dst + 2, static_cast<unsigned long long>(i2));
            }
        }
    }
}

// One-pass fixed point: terms are v * 2^e with e = kFixOneBits - (binary exponent of the
// largest characteristic acceleration G M / h^2 of the batch's systems, h the root cell's
// half size), so the resolution is 2^-kFixOneBits of that acceleration and a term stays
// below 2^49 (2^13 terms per leaf fit in 63 bits) for corrections up to 2^18 times it.
constexpr int kFixOneBits = 30;

// Two-pass fixed point, the fallback when a term leaves the one-pass range: the scale of
// a leaf whose largest |component| has the float bits `ub` multiplies its values by
// 2^(kFixBits - e), e the binary exponent of that maximum, so each |term| < 2^(kFixBits + 1)
// and fewer than 2^21 terms per leaf fit in 63 bits; the resolution is 2^-kFixBits of the
// largest term (5e-13 relative), far below an FP32 ulp of the sum.  Returns 0 for a leaf
// without (nonzero) corrections.
constexpr int kFixBits = 41;
__device__ __forceinline__ int fix_exp(uint32_t ub) {
    return (ub == 0u) ? 0 : kFixBits - (ilogb(static_cast<double>(__uint_as_float(ub))));
}

// Two-pass sums: per-leaf sums of the per-pair corrections, exact in 64-bit fixed point
// at each leaf's own unit (integer addition is associative, so the sums do not depend on
// the order in which the persistent walk listed the pairs).
// A wave takes kSumChunks chunks of 64 pairs per round, all of their loads in flight
// before the first segmented sum (one chunk at a time waited on two dependent loads).
constexpr int kSumChunks = 4;
__global__ __launch_bounds__(256) void k_learn_sum(const int2* __restrict__ pairs, const float4* __restrict__ pv,
                                                   const uint32_t* __restrict__ unit, const DevScalars* __restrict__ sc,
                                                   uint32_t cap, long long* __restrict__ acc) {
    const int lane = lane_id();
    const uint32_t np = min(sc->learn_pairs, cap);
    const int64_t nchunks = (static_cast<int64_t>(np) + 63) / 64;
    const int64_t nw = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
    for (int64_t ch0 = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); ch0 < nchunks;
         ch0 += kSumChunks * nw) {
        int leaf[kSumChunks];
        float4 v[kSumChunks];
        uint32_t ub[kSumChunks];
#pragma unroll
        for (int k = 0; k < kSumChunks; ++k) {
            const int64_t q = (ch0 + k * nw) * 64 + lane;
            leaf[k] = (q < np) ? pairs[q].x : -1;
        }
#pragma unroll
        for (int k = 0; k < kSumChunks; ++k) {
            const int64_t q = (ch0 + k * nw) * 64 + lane;
            v[k] = (leaf[k] >= 0) ? pv[q] : make_float4(0.f, 0.f, 0.f, 0.f);
            ub[k] = (leaf[k] >= 0) ? unit[leaf[k]] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kSumChunks; ++k) {
            const int e = fix_exp(ub[k]);
            long long i0 = __double2ll_rn(ldexp(static_cast<double>(v[k].x), e));
            long long i1 = __double2ll_rn(ldexp(static_cast<double>(v[k].y), e));
            long long i2 = __double2ll_rn(ldexp(static_cast<double>(v[k].z), e));
            const int lf = leaf[k];
            const int prev = __shfl_up(lf, 1, 64);
            const bool head = lane == 0 || prev != lf;
            const uint64_t heads = ballot(head);
            const int seg = __popcll(heads & ((2ull << lane) - 1ull));
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const long long t0 = __shfl_up(i0, d, 64), t1 = __shfl_up(i1, d, 64), t2 = __shfl_up(i2, d, 64);
                const int ts = __shfl_up(seg, d, 64);
                if (lane >= d && ts == seg) {
                    i0 += t0;
                    i1 += t1;
                    i2 += t2;
                }
            }
            const int next_seg = __shfl_down(seg, 1, 64);
            const bool tail = lane == 63 || next_seg != seg;
            if (tail && lf >= 0) {
                unsigned long long* dst = reinterpret_cast<unsigned long long*>(acc + 4 * static_cast<int64_t>(lf));
                atomicAdd(dst + 0, static_cast<unsigned long long>(i0));
                atomicAdd(dst + 1, static_cast<unsigned long long>(i1));
                atomicAdd(dst + 2, static_cast<unsigned long long>(i2));
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_learn_zero(long long* __restrict__ acc, uint32_t* __restrict__ unit,
                                                    DevScalars* __restrict__ sc, const int32_t* __restrict__ sys_root,
                                                    const double* __restrict__ sys_geom,
                                                    const float4* __restrict__ nd_rec, int nsys) {
    if (blockIdx.x == 0) {
        // the one-pass unit: the largest G M / h^2 over the systems
        __shared__ float s_a[256];
        float a = 0.f;
        for (int k = threadIdx.x; k < nsys; k += blockDim.x) {
            const double h = sys_geom[4 * k + 3];
            const float gm = nd_rec[4 * sys_root[k]].w;
            const float ak = static_cast<float>(static_cast<double>(gm) / (h * h));
            if (ak > a) a = ak;  // a NaN (no mass, no extent) is skipped
        }
        s_a[threadIdx.x] = a;
        __syncthreads();
        for (int o = 128; o >= 1; o >>= 1) {
            if (threadIdx.x < o) s_a[threadIdx.x] = fmaxf(s_a[threadIdx.x], s_a[threadIdx.x + o]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const float am = s_a[0];
            sc->learn_fix_e = (am > 0.f && am < 3.0e38f) ? kFixOneBits - ilogbf(am) : 0;
        }
    }
    const int nn = sc->n_nodes;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        reinterpret_cast<longlong2*>(acc)[2 * i] = make_longlong2(0, 0);
        reinterpret_cast<longlong2*>(acc)[2 * i + 1] = make_longlong2(0, 0);
        unit[i] = 0u;
    }
}

// The walk's accelerations (by sorted position) plus the correction of the particle's
// leaf, written once to input order.
template <bool TWO_PASS>
__global__ void k_learn_apply(const long long* __restrict__ acc, const uint32_t* __restrict__ unit,
                              const DevScalars* __restrict__ sc, const int32_t* __restrict__ leaf_of,
                              const uint32_t* __restrict__ perm, const float4* __restrict__ acc_sorted, int n,
                              float* __restrict__ ax, float* __restrict__ ay, float* __restrict__ az) {
    const int e1 = sc->learn_fix_e;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const float4 a = acc_sorted[p];
        const int lf = leaf_of[p];
        const int e = TWO_PASS ? fix_exp(unit[lf]) : e1;
        const longlong2 c01 = reinterpret_cast<const longlong2*>(acc)[2 * lf];
        const long long c2 = acc[4 * static_cast<int64_t>(lf) + 2];
        const uint32_t o = perm[p];
        ax[o] = a.x + static_cast<float>(ldexp(static_cast<double>(c01.x), -e));
        ay[o] = a.y + static_cast<float>(ldexp(static_cast<double>(c01.y), -e));
        az[o] = a.z + static_cast<float>(ldexp(static_cast<double>(c2), -e));
    }
}

// ---------------------------------------------------------------------------- samples
__device__ __forceinline__ double shfl_xor_f64(double v, int m) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__shfl_xor(hi, m, 64), __shfl_xor(lo, m, 64));
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += shfl_xor_f64(v, o);
    return v;
}

struct SampleArgs {
    const int2* pairs;
    int64_t npairs;
    int64_t nnew;
    const int32_t* nd_start;
    const int32_t* nd_count;
    const float4* nd_rec;  // 4 per node: com, quad, (Qyz, Qzz, first particle), aux
    const float4* nd_mom;
    const float4* spos;
    const uint32_t* perm;
    const float* amag2;
    double eps2;
    float* feat;  // 16 per sample, at the destination offset
    float* tgt;   // 4 per sample
};

// One wave per sample: pair k*npairs/nnew of the listed pairs.  The exact field of the
// cell on each member of the group is an FP64 sum over the cell's particles (self
// excluded); the multipole field is the walk's monopole + quadrupole formula evaluated
// in FP64 on the same FP32 moments.
__global__ __launch_bounds__(256) void k_learn_samples(SampleArgs A) {
    const int lane = lane_id();
    const int64_t wid = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    for (int64_t k = wid; k < A.nnew; k += nw) {
        const int64_t q = (A.nnew >= A.npairs) ? k : (k * A.npairs) / A.nnew;
        const int2 pr = A.pairs[q];
        const int gst = A.nd_start[pr.x], gcn = A.nd_count[pr.x];
        const int vst = A.nd_start[pr.y], vcn = A.nd_count[pr.y];
        const float4 cm = A.nd_rec[4 * pr.y];
        const float4 qd = A.nd_rec[4 * pr.y + 1];
        const float4 au = A.nd_rec[4 * pr.y + 3];
        double rx = 0.0, ry = 0.0, rz = 0.0, wa = 0.0;
        for (int mI = 0; mI < gcn; ++mI) {
            const float4 pm = A.spos[gst + mI];
            double ex = 0.0, ey = 0.0, ez = 0.0;
            for (int j = vst + lane; j < vst + vcn; j += 64) {
                if (j == gst + mI) continue;
                const float4 ps = A.spos[j];
                const double dx = static_cast<double>(ps.x) - pm.x, dy = static_cast<double>(ps.y) - pm.y,
                             dz = static_cast<double>(ps.z) - pm.z;
                const double r2 = dx * dx + dy * dy + dz * dz + A.eps2;
                const double inv = 1.0 / (r2 * sqrt(r2));
                const double f = static_cast<double>(ps.w) * inv;
                ex += f * dx;
                ey += f * dy;
                ez += f * dz;
            }
            ex = wave_sum_f64(ex);
            ey = wave_sum_f64(ey);
            ez = wave_sum_f64(ez);
            // monopole (softened) + quadrupole, as the walk evaluates an accepted cell
            const double dx = static_cast<double>(cm.x) - pm.x, dy = static_cast<double>(cm.y) - pm.y,
                         dz = static_cast<double>(cm.z) - pm.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double inv = 1.0 / sqrt(d2 + A.eps2);
            const double f = static_cast<double>(cm.w) * inv * inv * inv;
            const double qxx = qd.x, qyy = qd.y, qxy = qd.z, qxz = qd.w, qyz = au.x, qzz = -(qxx + qyy);
            const double qdx = qxx * dx + qxy * dy + qxz * dz;
            const double qdy = qxy * dx + qyy * dy + qyz * dz;
            const double qdz = qxz * dx + qyz * dy + qzz * dz;
            const double qd2 = dx * qdx + dy * qdy + dz * qdz;
            const double ir2 = 1.0 / d2, ir5 = ir2 * ir2 * sqrt(ir2);
            const double c2 = 2.5 * qd2 * ir5 * ir2;
            const double mx = f * dx + c2 * dx - ir5 * qdx;
            const double my = f * dy + c2 * dy - ir5 * qdy;
            const double mz = f * dz + c2 * dz - ir5 * qdz;
            rx += ex - mx;
            ry += ey - my;
            rz += ez - mz;
            const double a2 = A.amag2[A.perm[gst + mI]];
            wa += (a2 > 0.0) ? 1.0 / a2 : 0.0;
        }
        const float4 lc = A.nd_rec[4 * pr.x];
        const float4 tg = make_float4(lc.x, lc.y, lc.z, A.nd_rec[4 * pr.x + 3].z);
        float f[kF], scale;
        pair_features(tg, cm, au.y, A.nd_mom + 4 * pr.y, static_cast<float>(sqrt(A.eps2)), f, scale);
        const double inv_n = gcn > 0 ? 1.0 / gcn : 0.0;
        // entries 13 and 14 carry the pair's node ids (int32 bits) for inspection
        if (lane < 16)
            A.feat[16 * k + lane] = (lane < kF)     ? f[lane]
                                    : (lane == 13) ? __int_as_float(pr.x)
                                    : (lane == 14) ? __int_as_float(pr.y)
                                                   : 0.f;
        if (lane == 0) {
            const double is = 1.0 / scale;
            A.tgt[4 * k + 0] = static_cast<float>(rx * inv_n * is);
            A.tgt[4 * k + 1] = static_cast<float>(ry * inv_n * is);
            A.tgt[4 * k + 2] = static_cast<float>(rz * inv_n * is);
            A.tgt[4 * k + 3] = static_cast<float>(scale * sqrt(wa));
        }
    }
}

__global__ void k_learn_amag2(const float* __restrict__ ax, const float* __restrict__ ay,
                              const float* __restrict__ az, int n, float* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i];
}

// Pool statistics: per feature sum and sum of squares (FP64), and the weight sum.
__global__ __launch_bounds__(256) void k_learn_stats(const float* __restrict__ feat, const float* __restrict__ tgt,
                                                     int64_t n, double* __restrict__ out) {
    __shared__ double s[256][2];
    const int k = threadIdx.x & 15;   // feature of this thread
    const int r = threadIdx.x >> 4;   // row offset
    double a = 0.0, b = 0.0, w = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 16 + r; i < n; i += static_cast<int64_t>(gridDim.x) * 16) {
        const double v = feat[16 * i + k];
        a += v;
        b += v * v;
        if (k == 0) w += tgt[4 * i + 3];
    }
    s[threadIdx.x][0] = a;
    s[threadIdx.x][1] = b;
    __syncthreads();
    if (threadIdx.x < 16) {
        double sa = 0.0, sb = 0.0;
        for (int j = threadIdx.x; j < 256; j += 16) {
            sa += s[j][0];
            sb += s[j][1];
        }
        atomicAdd(out + threadIdx.x, sa);
        atomicAdd(out + 16 + threadIdx.x, sb);
    }
    __syncthreads();
    s[threadIdx.x][0] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sw = 0.0;
        for (int j = 0; j < 256; ++j) sw += s[j][0];
        atomicAdd(out + 32, sw);
    }
}

// ---------------------------------------------------------------------------- training
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct TrainArgs {
    const float* feat;  // 16 per sample
    const float* tgt;   // 4 per sample
    int64_t n_src;      // samples available
    int use_hash;       // 1: sample i of block b = hash(seed, step, b*64+i) mod n_src; 0: row b*64+i
    uint64_t seed;
    int step;
    float inv_wmean;    // loss weight multiplier
    float inv_n;        // 1 / samples in the batch
    const uint16_t* w16;
    const float* w32;   // full blob
    float* grad_part;   // per block, kNTrain
    float* loss_part;   // per block
    float* out;         // optional predictions (n x 3, direct rows only)
};

// One wave per block, 64 samples: forward with the pre-activations kept in LDS, then the
// backward pass; weight gradients are matrix-core products over the block's samples.
template <int ACT>
__global__ __launch_bounds__(64) void k_learn_fwdbwd(TrainArgs A) {
    __shared__ __attribute__((aligned(16))) uint16_t s_x[64][16];
    __shared__ __attribute__((aligned(16))) float s_z[3][64][kH];
    __shared__ __attribute__((aligned(16))) uint16_t s_d[64][kH];
    __shared__ __attribute__((aligned(16))) float s_y[64][4];
    const int lane = lane_id();
    const int g = lane >> 4, c = lane & 15;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * 64 + lane;
    int64_t src = row;
    bool valid = true;
    if (A.use_hash) {
        src = static_cast<int64_t>(mix64(A.seed ^ mix64(static_cast<uint64_t>(A.step) * 0x100000001b3ull +
                                                        static_cast<uint64_t>(row))) %
                                   static_cast<uint64_t>(A.n_src));
    } else {
        valid = row < A.n_src;
        if (!valid) src = 0;
    }
    const float* mu = A.w32;
    const float* isd = A.w32 + 16;
    const float* b32 = A.w32 + kNorm;
    // inputs
    const float4* fp = reinterpret_cast<const float4*>(A.feat + 16 * src);
    float fin[16];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float4 v = fp[k];
        fin[4 * k] = v.x;
        fin[4 * k + 1] = v.y;
        fin[4 * k + 2] = v.z;
        fin[4 * k + 3] = v.w;
    }
    const float4 tg = reinterpret_cast<const float4*>(A.tgt)[src];
    const float wgt = valid ? tg.w * A.inv_wmean : 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) s_x[lane][k] = (k < kF) ? bf16_rne((fin[k] - mu[k]) * isd[k]) : 0;
    __syncthreads();
    // forward
    {
        Frags F;
        load_frags(F, A.w16, b32, lane);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x4 x = *reinterpret_cast<const bf16x4*>(&s_x[16 * t + c][4 * g]);
            const f32x4 y = mlp_tile<true>(F, x, &s_z[0][0][0], 16 * t, lane, ActDirect<ACT>{});
            if (g == 0) {
                s_y[16 * t + c][0] = y[0];
                s_y[16 * t + c][1] = y[1];
                s_y[16 * t + c][2] = y[2];
            }
        }
    }
    __syncthreads();
    // loss and dL/dy (lane = sample)
    float dy0, dy1, dy2;
    {
        const float e0 = s_y[lane][0] - tg.x, e1 = s_y[lane][1] - tg.y, e2 = s_y[lane][2] - tg.z;
        const float l = wgt * (e0 * e0 + e1 * e1 + e2 * e2);
        const float ls = wave_sum(l);
        if (lane == 0) A.loss_part[blockIdx.x] = ls;
        const float k2 = 2.f * wgt * A.inv_n;
        dy0 = k2 * e0;
        dy1 = k2 * e1;
        dy2 = k2 * e2;
        if (A.out && valid) {
            A.out[3 * row + 0] = s_y[lane][0];
            A.out[3 * row + 1] = s_y[lane][1];
            A.out[3 * row + 2] = s_y[lane][2];
        }
    }
    __syncthreads();
    s_y[lane][0] = dy0;
    s_y[lane][1] = dy1;
    s_y[lane][2] = dy2;
    __syncthreads();
    float* gp = A.grad_part + static_cast<int64_t>(blockIdx.x) * kNTrain;
    // output layer: dW4[i][o] = sum_n h3[n][i] dy[n][o] (lane = i), db4, then dZ3 (lane = n)
    {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        for (int n = 0; n < 64; ++n) {
            const float h = bf16_f32(act_of<ACT>(s_z[2][n][lane]));
            a0 += h * s_y[n][0];
            a1 += h * s_y[n][1];
            a2 += h * s_y[n][2];
        }
        gp[kOffW4 + 3 * lane + 0] = a0;
        gp[kOffW4 + 3 * lane + 1] = a1;
        gp[kOffW4 + 3 * lane + 2] = a2;
        if (lane < 3) {
            float s = 0.f;
            for (int n = 0; n < 64; ++n) s += s_y[n][lane];
            gp[kOffB4 + lane] = s;
        }
        const uint16_t* w4 = A.w16 + kOffW4;  // the BF16 W4[i][o] the forward used
        for (int i = 0; i < kH; ++i) {
            const float dh = dy0 * bf16_f32(w4[3 * i]) + dy1 * bf16_f32(w4[3 * i + 1]) + dy2 * bf16_f32(w4[3 * i + 2]);
            s_d[lane][i] = bf16_rne(dh * dact<ACT>(s_z[2][lane][i]));
        }
    }
    __syncthreads();
    // hidden layers, L = 3, 2: dW_L from (h_{L-1}, dZ_L), then dZ_{L-1}
    // dZ3 lives in s_d; dZ2 is written over z3 (s_z[2], dead); dZ1 over s_d.
    uint16_t* d3 = &s_d[0][0];
    uint16_t* d2 = reinterpret_cast<uint16_t*>(&s_z[2][0][0]);
#pragma unroll
    for (int L = 3; L >= 2; --L) {
        const uint16_t* dcur = (L == 3) ? d3 : d2;
        uint16_t* dnext = (L == 3) ? d2 : d3;
        const float* zin = &s_z[L - 2][0][0];  // pre-activations of the layer's input
        const int offW = (L == 3) ? kOffW3 : kOffW2;
        const int offB = (L == 3) ? kOffB3 : kOffB2;
        // dW[i][h] = sum_n h_in[n][i] dZ[n][h]: D rows i, cols h, K = samples
        {
            bf16x4 Ain[4][4];  // [ib][nb]: A[i = 16 ib + c][n = 16 nb + 4g + j]
            bf16x4 Bd[4][4];   // [nb][hb]: B[n = 16 nb + 4g + j][h = 16 hb + c]
#pragma unroll
            for (int ib = 0; ib < 4; ++ib)
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    bf16x4 v;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        v[j] = static_cast<short>(act_of<ACT>(zin[(16 * nb + 4 * g + j) * kH + 16 * ib + c]));
                    Ain[ib][nb] = v;
                }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
#pragma unroll
                for (int hb = 0; hb < 4; ++hb) {
                    bf16x4 v;
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = static_cast<short>(dcur[(16 * nb + 4 * g + j) * kH + 16 * hb + c]);
                    Bd[nb][hb] = v;
                }
#pragma unroll
            for (int ib = 0; ib < 4; ++ib)
#pragma unroll
                for (int hb = 0; hb < 4; ++hb) {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int nb = 0; nb < 4; ++nb) acc = mfma16(Ain[ib][nb], Bd[nb][hb], acc);
#pragma unroll
                    for (int j = 0; j < 4; ++j) gp[offW + (16 * ib + 4 * g + j) * kH + 16 * hb + c] = acc[j];
                }
        }
        // db[h] = sum_n dZ[n][h] (lane = h)
        {
            float s = 0.f;
            for (int n = 0; n < 64; ++n) s += bf16_f32(dcur[n * kH + lane]);
            gp[offB + lane] = s;
        }
        // dH_in^T[i][n] = sum_h W[i][h] dZ^T[h][n]; dZ_in = dH_in * act'(z_in)
        {
            f32x4 acc[4][4];  // [ib][nb]
#pragma unroll
            for (int ib = 0; ib < 4; ++ib)
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) acc[ib][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                bf16x4 Bt[4];
#pragma unroll
                for (int nb = 0; nb < 4; ++nb)
                    Bt[nb] = *reinterpret_cast<const bf16x4*>(dcur + (16 * nb + c) * kH + 16 * kb + 4 * g);
#pragma unroll
                for (int ib = 0; ib < 4; ++ib) {
                    const bf16x4 a =
                        *reinterpret_cast<const bf16x4*>(A.w16 + offW + (16 * ib + c) * kH + 16 * kb + 4 * g);
#pragma unroll
                    for (int nb = 0; nb < 4; ++nb) acc[ib][nb] = mfma16(a, Bt[nb], acc[ib][nb]);
                }
            }
            __syncthreads();  // dcur and z_out fully read before dnext (may alias z_out) is written
#pragma unroll
            for (int ib = 0; ib < 4; ++ib)
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    const int n = 16 * nb + c;
                    bf16x4 v;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int i = 16 * ib + 4 * g + j;
                        v[j] = static_cast<short>(bf16_rne(acc[ib][nb][j] * dact<ACT>(zin[n * kH + i])));
                    }
                    *reinterpret_cast<bf16x4*>(dnext + n * kH + 16 * ib + 4 * g) = v;
                }
        }
        __syncthreads();
    }
    // first layer: dW1[f][h] = sum_n x[n][f] dZ1[n][h]; db1
    {
        const uint16_t* d1 = d3;
        bf16x4 Ax[4];  // [nb]: A[f = c][n = 16 nb + 4g + j]
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            bf16x4 v;
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = static_cast<short>(s_x[16 * nb + 4 * g + j][c]);
            Ax[nb] = v;
        }
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                bf16x4 v;
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = static_cast<short>(d1[(16 * nb + 4 * g + j) * kH + 16 * hb + c]);
                acc = mfma16(Ax[nb], v, acc);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int f = 4 * g + j;
                if (f < kF) gp[kOffW1 + f * kH + 16 * hb + c] = acc[j];
            }
        }
        float s = 0.f;
        for (int n = 0; n < 64; ++n) s += bf16_f32(d1[n * kH + lane]);
        gp[kOffB1 + lane] = s;
    }
}

// Sum of the block partials; Adam update of the FP32 master and its BF16 copy.
__global__ __launch_bounds__(256) void k_learn_adam(const float* __restrict__ grad_part, int nblocks,
                                                    const float* __restrict__ loss_part, float* __restrict__ loss_slot,
                                                    float* __restrict__ w32, uint16_t* __restrict__ w16,
                                                    float* __restrict__ m, float* __restrict__ v, float lr, float bc1,
                                                    float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_slot) {
        float s = 0.f;
        for (int b = 0; b < nblocks; ++b) s += loss_part[b];
        *loss_slot = s;
    }
    if (p >= kNTrain) return;
    float gsum = 0.f;
    for (int b = 0; b < nblocks; ++b) gsum += grad_part[static_cast<int64_t>(b) * kNTrain + p];
    const float b1 = 0.9f, b2 = 0.999f;
    const float mm = b1 * m[p] + (1.f - b1) * gsum;
    const float vv = b2 * v[p] + (1.f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    const float w = w32[kNorm + p] - lr * (mm * bc1) / (sqrtf(vv * bc2) + 1e-8f);
    w32[kNorm + p] = w;
    w16[p] = bf16_rne(w);
}

__global__ __launch_bounds__(256) void k_learn_reduce(const float* __restrict__ grad_part, int nblocks,
                                                      const float* __restrict__ loss_part, float* __restrict__ grad,
                                                      float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float s = 0.f;
        for (int b = 0; b < nblocks; ++b) s += loss_part[b];
        *loss = s;
    }
    if (p >= kNTrain) return;
    float gsum = 0.f;
    for (int b = 0; b < nblocks; ++b) gsum += grad_part[static_cast<int64_t>(b) * kNTrain + p];
    grad[p] = gsum;
}

// ---------------------------------------------------------------------------- launchers
namespace {
template <typename T>
int lalloc(nbg_ctx* c, T*& p, size_t count) {
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
}  // namespace

// Allocates whatever of the learned buffers is missing (a failed earlier attempt leaves
// some of them unset), so that a learned evaluation never runs without all of them.
int learn_alloc(nbg_ctx* c) {
    const size_t C = static_cast<size_t>(c->cap_nodes);
    if (!c->nd_mom)
        if (int r = lalloc(c, c->nd_mom, 4 * C)) return r;
    if (!c->lacc)
        if (int r = lalloc(c, c->lacc, 4 * C)) return r;
    if (!c->lunit)
        if (int r = lalloc(c, c->lunit, C)) return r;
    if (!c->amag2)
        if (int r = lalloc(c, c->amag2, static_cast<size_t>(c->cap_n))) return r;
    if (!c->lpairs || !c->lpv) return learn_grow_pairs(c, std::max<int64_t>(int64_t{1} << 20, 32 * c->cap_n));
    return NBG_OK;
}

int learn_grow_pairs(nbg_ctx* c, int64_t need) {
    if (need <= c->lpairs_cap && c->lpairs && c->lpv) return NBG_OK;
    const int64_t cap = std::min<int64_t>(need + need / 4 + 1024, int64_t{0xffffffffll});
    int r = lalloc(c, c->lpairs, static_cast<size_t>(cap));
    if (r == NBG_OK) r = lalloc(c, c->lpv, static_cast<size_t>(cap));
    if (r != NBG_OK) {
        c->lpairs_cap = 0;
        return r;
    }
    c->lpairs_cap = cap;
    return NBG_OK;
}

void launch_learn_zero(nbg_ctx* c) {
    hipLaunchKernelGGL(k_learn_zero, dim3(1024), dim3(256), 0, c->stream, c->lacc, c->lunit, c->dsc, c->sys_root,
                       c->sys_geom, c->nd_rec, c->nsys);
}

void launch_learn_correct(nbg_ctx* c, const nbg_params& p) {
    InferArgs A;
    A.pairs = c->lpairs;
    A.sc = c->dsc;
    A.scw = c->dsc;
    A.cap = static_cast<uint32_t>(c->lpairs_cap);
    A.nd_rec = c->nd_rec;
    A.nd_mom = c->nd_mom;
    A.eps = static_cast<float>(std::sqrt(p.eps2));
    A.w16 = c->lw16;
    A.w32 = c->lw32;
    A.pv = c->lpv;
    A.unit = c->lunit;
    A.acc = c->lacc;
    const dim3 blk(64 * kInferWaves), grid(kInferCUs);
    const bool relu = c->learn_act == kActRelu;
    if (c->learn_two_pass) {
        if (relu)
            hipLaunchKernelGGL((k_learn_infer<true, kActRelu>), grid, blk, 0, c->stream, A);
        else
            hipLaunchKernelGGL((k_learn_infer<true, kActSilu>), grid, blk, 0, c->stream, A);
        hipLaunchKernelGGL(k_learn_sum, dim3(2048), dim3(256), 0, c->stream, c->lpairs, c->lpv, c->lunit, c->dsc,
                           static_cast<uint32_t>(c->lpairs_cap), c->lacc);
    } else {
        if (relu)
            hipLaunchKernelGGL((k_learn_infer<false, kActRelu>), grid, blk, 0, c->stream, A);
        else
            hipLaunchKernelGGL((k_learn_infer<false, kActSilu>), grid, blk, 0, c->stream, A);
    }
    launch_learn_apply(c);
}

void launch_learn_apply(nbg_ctx* c) {
    if (c->learn_two_pass)
        hipLaunchKernelGGL(k_learn_apply<true>, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->lacc,
                           c->lunit, c->dsc, c->leaf_of, c->perm, c->stage4, c->n, c->ax, c->ay, c->az);
    else
        hipLaunchKernelGGL(k_learn_apply<false>, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->lacc,
                           c->lunit, c->dsc, c->leaf_of, c->perm, c->stage4, c->n, c->ax, c->ay, c->az);
}

void launch_learn_amag2(nbg_ctx* c) {
    hipLaunchKernelGGL(k_learn_amag2, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->ax, c->ay, c->az, c->n,
                       c->amag2);
}

void launch_learn_samples(nbg_ctx* c, const nbg_params& p, int64_t npairs, int64_t nnew, int64_t dst) {
    SampleArgs A;
    A.pairs = c->lpairs;
    A.npairs = npairs;
    A.nnew = nnew;
    A.nd_start = c->nd_start;
    A.nd_count = c->nd_count;
    A.nd_rec = c->nd_rec;
    A.nd_mom = c->nd_mom;
    A.spos = c->spos;
    A.perm = c->perm;
    A.amag2 = c->amag2;
    A.eps2 = p.eps2;
    A.feat = c->pool_feat + 16 * dst;
    A.tgt = c->pool_tgt + 4 * dst;
    hipLaunchKernelGGL(k_learn_samples, dim3(4096), dim3(256), 0, c->stream, A);
}

void launch_learn_stats(nbg_ctx* c, int64_t count, double* out) {
    hipLaunchKernelGGL(k_learn_stats, dim3(1024), dim3(256), 0, c->stream, c->pool_feat, c->pool_tgt, count, out);
}

void launch_learn_fwdbwd(nbg_ctx* c, const float* feat, const float* tgt, int64_t n_src, int nblocks, uint64_t seed,
                         int step, int use_hash, float inv_wmean, float inv_n, float* out) {
    TrainArgs A;
    A.feat = feat;
    A.tgt = tgt;
    A.n_src = n_src;
    A.use_hash = use_hash;
    A.seed = seed;
    A.step = step;
    A.inv_wmean = inv_wmean;
    A.inv_n = inv_n;
    A.w16 = c->lw16;
    A.w32 = c->lw32;
    A.grad_part = c->grad_part;
    A.loss_part = c->loss_part;
    A.out = out;
    if (c->learn_act == kActRelu)
        hipLaunchKernelGGL(k_learn_fwdbwd<kActRelu>, dim3(nblocks), dim3(64), 0, c->stream, A);
    else
        hipLaunchKernelGGL(k_learn_fwdbwd<kActSilu>, dim3(nblocks), dim3(64), 0, c->stream, A);
}

void launch_learn_adam(nbg_ctx* c, int nblocks, float lr, int64_t t, float* loss_slot) {
    const double b1 = 0.9, b2 = 0.999;
    const float bc1 = static_cast<float>(1.0 / (1.0 - std::pow(b1, static_cast<double>(t))));
    const float bc2 = static_cast<float>(1.0 / (1.0 - std::pow(b2, static_cast<double>(t))));
    hipLaunchKernelGGL(k_learn_adam, dim3(grid_for(kNTrain, 256)), dim3(256), 0, c->stream, c->grad_part, nblocks,
                       c->loss_part, loss_slot, c->lw32, c->lw16, c->adam_m, c->adam_v, lr, bc1, bc2);
}

void launch_learn_reduce(nbg_ctx* c, int nblocks, float* grad_out, float* loss_out) {
    hipLaunchKernelGGL(k_learn_reduce, dim3(grid_for(kNTrain, 256)), dim3(256), 0, c->stream, c->grad_part, nblocks,
                       c->loss_part, grad_out, loss_out);
}

}  // namespace nbg
