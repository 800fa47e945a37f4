// mlp_device.h -- the learned Newton seed on the device.
//
// A 51 -> 64 -> 64 -> 64 -> 3 network (tanh-form GELU) predicts the
// converged (u, v, t) of newton_refine (SurfacePatch.hpp:593-660) for a
// (ray, patch) pair.  The reference has no such seed; its seed is the centroid
// of the closest stage-1 triangle (SurfacePatch.hpp:272-303, :499-510).
//
// Matrix work runs on the gfx950 matrix cores, bf16 operands and fp32
// accumulation: v_mfma_f32_16x16x32_bf16 for every 64-wide contraction (two
// 16-wide K slices per instruction), v_mfma_f32_16x16x16_bf16 where only one
// slice exists (layer 1's six ray columns, the output layer's transpose).  Every layer is computed transposed,
// Z^T = W X^T, with the weights as the A operand and the activations as the B
// operand.  The instruction's D layout (lane l holds rows 4(l/16)+i, column
// l%16) is then exactly the B layout the next layer needs (lane l holds rows
// 4(l/16)+i of column l%16), so a whole forward pass stays in registers: lane
// l owns sample l%16 of each 16-sample tile and features 4(l/16)..+3 of each
// 16-feature tile.
//
// Weights are kept in "fragment order": fragment f, lane l is the 4 bf16
// values lane l feeds the instruction, so a wave loads a fragment with one
// coalesced 512-byte read.
#pragma once
#include "device_common.h"

namespace mlp {

constexpr int kIn = 51, kRay = 6, kPatchF = 45, kH = 64, kOut = 3, kOutPad = 16;
// parameter layout (fp32 master copy); the pads are zero and receive no update
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kH;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH * kH;
constexpr int kOffW4 = kOffB3 + kH;
constexpr int kOffB4 = kOffW4 + kOutPad * kH;
constexpr int kNParams = kOffB4 + kOutPad;
static_assert(kNParams == BBRT_MLP_NPARAMS, "parameter layout");

// fragment-ordered bf16 weights, [fragment][64 lanes] of 4 bf16 each
constexpr int kFragFwd1 = 0;    // W1:  (ot, kt) -> ot*4 + kt,  16 fragments
constexpr int kFragFwd2 = 16;   // W2
constexpr int kFragFwd3 = 32;   // W3
constexpr int kFragFwd4 = 48;   // W4:  kt, 4 fragments
constexpr int kFragBwd2 = 52;   // W2^T: (it, ot) -> it*4 + ot, 16 fragments
constexpr int kFragBwd3 = 68;   // W3^T
constexpr int kFragBwd4 = 84;   // W4^T: it, 4 fragments
constexpr int kFragRay1 = 88;   // W1 columns 0..5 (the ray features), zero elsewhere: ot, 4 fragments
constexpr int kNFrags = 92;

constexpr int kFrameDoubles = 24;   // per patch: centre[3], e1[3], e2[3], n[3], scale, pad[3], a_u[3], a_v[3], pad[2]
constexpr double kPlaneClamp = 1.5;  // tangent-plane estimate clamped to [-1.5, 1.5]
constexpr int kPatchFeatPad = 48;   // per patch: 45 bf16 features + 3 zero

// The seed the network predicts is accepted only if Newton from it converges
// (converged-mode classification and polish) to a root within this many
// parameter units of the stage-1 triangle seed: two coarse cells.
constexpr double kSeedWindow = 2.0 * 2.0 / 99.0;

}  // namespace mlp

namespace {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// float -> bfloat16, round to nearest even (inputs are finite)
__device__ __forceinline__ uint32_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf2f(uint32_t b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ u32x2_t pack4(float a, float b, float c, float d) {
    u32x2_t r;
    r.x = f2bf(a) | (f2bf(b) << 16);
    r.y = f2bf(c) | (f2bf(d) << 16);
    return r;
}

// D = A B + C for 16x16 tiles: A 16x16 (lane l: row l%16, columns 4(l/16)..+3),
// B 16x16 (lane l: rows 4(l/16)..+3, column l%16), C/D (lane l: rows 4(l/16)+i, column l%16).
__device__ __forceinline__ f32x4_t mfma16(u32x2_t a, u32x2_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(bf16x4_t, a), __builtin_bit_cast(bf16x4_t, b),
                                                     c, 0, 0, 0);
}

// D = A B + C contracting over 32: two 16-wide K slices (a0, b0) and (a1, b1)
// in mfma16's lane layout, issued as one gfx950 16x16x32 instruction.  Lane
// group l/16 of A and of B carries the same eight K indices (4 of slice 0,
// then 4 of slice 1), so the product is the sum over both slices whatever
// order the instruction walks K in.
__device__ __forceinline__ f32x4_t mfma32(u32x2_t a0, u32x2_t a1, u32x2_t b0, u32x2_t b1, f32x4_t c) {
    const u32x4_t a = { a0.x, a0.y, a1.x, a1.y }, b = { b0.x, b0.y, b1.x, b1.y };
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                   0, 0, 0);
}

// GELU, tanh form: 0.5 z (1 + tanh(k0 (z + k1 z^3))), and its derivative.
__device__ __forceinline__ float gelu_f(float z, float& gp) {
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float a = k0 * (z + k1 * z * z * z);
    const float e = __expf(2.0f * a);
    const float T = 1.0f - 2.0f / (e + 1.0f);
    gp = 0.5f * (1.0f + T) + 0.5f * z * (1.0f - T * T) * k0 * (1.0f + 3.0f * k1 * z * z);
    return 0.5f * z * (1.0f + T);
}

// One hidden layer over a wave's 64 samples: out = gelu(W in + b).
// in[kt][st] / out[ot][st]: B fragments (feature tile, sample tile).
// Wf: 16 fragments in (ot*4 + kt) order.  gp (optional): GELU' in D order,
// [ot*4 + st][64 lanes].
template <bool STORE_GP>
__device__ __forceinline__ void hidden_layer(const u32x2_t* Wf, const float* b, const u32x2_t (&in)[4][4],
                                             u32x2_t (&out)[4][4], f32x4_t* gp, int lane) {
    const int q = lane >> 4;
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        u32x2_t w[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) w[kt] = Wf[(ot * 4 + kt) * 64 + lane];
        const f32x4_t bias = { b[ot * 16 + 4 * q], b[ot * 16 + 4 * q + 1], b[ot * 16 + 4 * q + 2],
                               b[ot * 16 + 4 * q + 3] };
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            f32x4_t acc = bias;
#pragma unroll
            for (int kp = 0; kp < 2; ++kp) acc = mfma32(w[2 * kp], w[2 * kp + 1], in[2 * kp][st], in[2 * kp + 1][st], acc);
            float g0, g1, g2, g3;
            const float h0 = gelu_f(acc[0], g0), h1 = gelu_f(acc[1], g1), h2 = gelu_f(acc[2], g2),
                        h3 = gelu_f(acc[3], g3);
            out[ot][st] = pack4(h0, h1, h2, h3);
            if constexpr (STORE_GP) gp[(ot * 4 + st) * 64 + lane] = f32x4_t{ g0, g1, g2, g3 };
        }
    }
}

// Output layer: y[st] (D order; lanes 0..15 hold outputs 0..3 of sample st*16 + lane).
__device__ __forceinline__ void out_layer(const u32x2_t* Wf4, const float* b4, const u32x2_t (&in)[4][4],
                                          f32x4_t (&y)[4], int lane) {
    const int q = lane >> 4;
    u32x2_t w[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) w[kt] = Wf4[kt * 64 + lane];
    const f32x4_t bias = { b4[4 * q], b4[4 * q + 1], b4[4 * q + 2], b4[4 * q + 3] };
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        f32x4_t acc = bias;
#pragma unroll
        for (int kp = 0; kp < 2; ++kp) acc = mfma32(w[2 * kp], w[2 * kp + 1], in[2 * kp][st], in[2 * kp + 1][st], acc);
        y[st] = acc;
    }
}

// The network as the bounce kernel holds it.
struct MlpDev {
    const double* frame;     // [n_patches][kFrameDoubles]
    const float* h1pre;      // [n_patches][64]: b1 + W1[:, 6:51] f_patch, fp32
    const u32x2_t* wfrag;    // [kNFrags][64]
    const float* params;     // fp32 master (biases)
};

// What one workgroup stages in LDS for inference: the ray fragments of W1,
// W2, W3 and W4 in fragment order, then b2, b3, b4.
struct MlpLds {
    u32x2_t w[(4 + 16 + 16 + 4) * 64];
    float b[64 + 64 + 16];
};

__device__ __forceinline__ void mlp_stage_lds(const MlpDev& M, MlpLds& L, int tid, int nthreads) {
    for (int g = tid; g < 40 * 64; g += nthreads) {
        const int f = g >> 6, l = g & 63;
        const int src = f < 4 ? mlp::kFragRay1 + f : (f < 20 ? mlp::kFragFwd2 + f - 4 : (f < 36 ? mlp::kFragFwd3 + f - 20
                                                                                              : mlp::kFragFwd4 + f - 36));
        L.w[g] = M.wfrag[src * 64 + l];
    }
    for (int g = tid; g < 144; g += nthreads)
        L.b[g] = g < 64 ? M.params[mlp::kOffB2 + g] : (g < 128 ? M.params[mlp::kOffB3 + g - 64] : M.params[mlp::kOffB4 + g - 128]);
}

// Ray features of a (ray, patch) pair: the ray in the patch frame, its origin
// moved to the point closest to the frame centre (parameter tc); and the
// tangent-plane estimate (ua, va) of the hit's parameters, which the network
// corrects.  FP64, then rounded to fp32 -- the same operations in the same
// order as the oracle.
struct RayFeat {
    float r[6];
    double tc, s, ua, va;
};
__device__ __forceinline__ RayFeat ray_features(const double* __restrict__ F, const V3& o, const V3& d) {
    const V3 c = { F[0], F[1], F[2] }, e1 = { F[3], F[4], F[5] }, e2 = { F[6], F[7], F[8] }, n = { F[9], F[10], F[11] };
    RayFeat f;
    f.s = F[12];
    const V3 cmo = sub3(c, o);
    f.tc = dot3(cmo, d);
    const V3 q = { o.x + f.tc * d.x - c.x, o.y + f.tc * d.y - c.y, o.z + f.tc * d.z - c.z };
    f.r[0] = (float)(dot3(q, e1) / f.s);
    f.r[1] = (float)(dot3(q, e2) / f.s);
    f.r[2] = (float)(dot3(q, n) / f.s);
    f.r[3] = (float)dot3(d, e1);
    f.r[4] = (float)dot3(d, e2);
    f.r[5] = (float)dot3(d, n);
    const double dn = dot3(d, n);
    const double tp = fabs(dn) > 1e-12 ? dot3(cmo, n) / dn : f.tc;
    const V3 pc = { o.x + tp * d.x - c.x, o.y + tp * d.y - c.y, o.z + tp * d.z - c.z };
    const V3 au = { F[16], F[17], F[18] }, av = { F[19], F[20], F[21] };
    f.ua = fmin(fmax(dot3(au, pc), -mlp::kPlaneClamp), mlp::kPlaneClamp);
    f.va = fmin(fmax(dot3(av, pc), -mlp::kPlaneClamp), mlp::kPlaneClamp);
    return f;
}

// Learned seeds (u0, v0, t0) = (ua + y0, va + y1, tc + s y2) for the lanes with
// `need`; every lane of the wave must call this (the matrix instructions use
// the whole wave).
__device__ __forceinline__ void mlp_seed_wave(const MlpDev& M, const MlpLds& L, bool need, int32_t patch, const V3& o,
                                              const V3& d, double& u0, double& v0, double& t0) {
    const int lane = __lane_id(), q = lane >> 4, j = lane & 15;
    uint32_t p01 = 0, p23 = 0, p45 = 0;
    RayFeat f{ { 0, 0, 0, 0, 0, 0 }, 0.0, 1.0, 0.0, 0.0 };
    if (need) {
        f = ray_features(M.frame + (size_t)patch * mlp::kFrameDoubles, o, d);
        p01 = f2bf(f.r[0]) | (f2bf(f.r[1]) << 16);
        p23 = f2bf(f.r[2]) | (f2bf(f.r[3]) << 16);
        p45 = f2bf(f.r[4]) | (f2bf(f.r[5]) << 16);
    }
    const int my_pid = need ? patch : -1;
    u32x2_t h1[4][4];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const int src = st * 16 + j;
        const uint32_t a = __shfl(p01, src), b = __shfl(p23, src), c = __shfl(p45, src);
        const int pid = __shfl(my_pid, src);
        u32x2_t x;
        x.x = q == 0 ? a : (q == 1 ? c : 0u);
        x.y = q == 0 ? b : 0u;
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            f32x4_t acc = { 0.0f, 0.0f, 0.0f, 0.0f };
            if (pid >= 0) {
                const float4 hp = *reinterpret_cast<const float4*>(M.h1pre + (size_t)pid * 64 + ot * 16 + 4 * q);
                acc = f32x4_t{ hp.x, hp.y, hp.z, hp.w };
            }
            acc = mfma16(L.w[ot * 64 + lane], x, acc);
            float g;
            h1[ot][st] = pack4(gelu_f(acc[0], g), gelu_f(acc[1], g), gelu_f(acc[2], g), gelu_f(acc[3], g));
        }
    }
    u32x2_t h2[4][4], h3[4][4];
    hidden_layer<false>(L.w + 4 * 64, L.b, h1, h2, nullptr, lane);
    hidden_layer<false>(L.w + 20 * 64, L.b + 64, h2, h3, nullptr, lane);
    f32x4_t y[4];
    out_layer(L.w + 36 * 64, L.b + 128, h3, y, lane);
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const float a = __shfl(y[st][0], j), b = __shfl(y[st][1], j), c = __shfl(y[st][2], j);
        if (q == st) { y0 = a; y1 = b; y2 = c; }
    }
    u0 = f.ua + (double)y0;
    v0 = f.va + (double)y1;
    t0 = f.tc + f.s * (double)y2;
}

}  // namespace
