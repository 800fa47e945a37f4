// bbrt_mlp.hip -- the learned Newton seed: pair generation, training and
// evaluation kernels, and their C ABI.
//
// Training data are (ray, patch) pairs that the method solves itself: a ray
// is aimed at a random point of a random patch from its front side, stage 1 +
// converged Newton (SurfacePatch.hpp:253-303, :593-660, as in the bounce
// kernel) find the root, and the pair keeps that root as its target.
//
// Training runs one wave per 64-sample tile: the forward pass stays in
// registers (see mlp_device.h), activations are also written to LDS
// feature-major for the weight gradients, which are bf16 matrix-core products
// over the tile's samples with fp32 accumulation.  Each tile writes its
// gradient; one thread per parameter sums the tiles in a fixed order (so a
// step is deterministic) and applies Adam to the fp32 master weights, which
// are then re-packed to bf16 fragments.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mlp_device.h"
#include "runtime.h"

using namespace mlp;

namespace {

constexpr int kFeatRow = 16;   // u32x2 per training sample: 64 bf16 (51 features + zero pad)

// ---------------------------------------------------------------------------
// Random pairs
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// draw k (0-based) of the SplitMix64 stream started at `seed`, as a double in [0, 1)
__device__ __forceinline__ double sm_uniform(uint64_t seed, uint64_t k) {
    return (double)(sm_mix(seed + 0x9E3779B97F4A7C15ull * (k + 1)) >> 11) * (1.0 / 9007199254740992.0);
}

// X (fields f0..f0+2) of patch k at (u, v) with the reference's basis and field sum.
__device__ V3 eval3(const DevScene& S, const PatchMeta& M, double u, double v, int f0) {
    const int im = M.imax, jm = M.jmax;
    const double* un = S.nodes + M.node_off;
    const double* uw = un + im;
    const double* vn = uw + im;
    const double* vw = vn + jm;
    const double* F = S.fields + M.field_off;
    const int nn = im * jm;
    double lu[bbrt::kMaxNodes1D], lv[bbrt::kMaxNodes1D];
    basis<0>(un, uw, im, u, lu);
    basis<0>(vn, vw, jm, v, lv);
    return { field<0>(lu, lv, F + f0 * nn, im, jm), field<0>(lu, lv, F + (f0 + 1) * nn, im, jm),
             field<0>(lu, lv, F + (f0 + 2) * nn, im, jm) };
}

struct Pair {
    int32_t patch;
    V3 o, d;
};

// Pair i: patch uniform, (u, v) uniform in [-1, 1]^2, direction uniform in
// solid angle over the front hemisphere with |cos| >= 0.02, origin 0.25..4
// patch scales back along the direction.
__device__ Pair draw_pair(const DevScene& S, const double* frame, int64_t n_patch, uint64_t seed, int64_t i) {
    const uint64_t k = (uint64_t)i * 8;
    const double r0 = sm_uniform(seed, k), r1 = sm_uniform(seed, k + 1), r2 = sm_uniform(seed, k + 2);
    const double r3 = sm_uniform(seed, k + 3), r4 = sm_uniform(seed, k + 4), r5 = sm_uniform(seed, k + 5);
    int64_t p = (int64_t)(r0 * (double)n_patch);
    if (p >= n_patch) p = n_patch - 1;
    const PatchMeta M = S.meta[p];
    const double u = -1.0 + 2.0 * r1, v = -1.0 + 2.0 * r2;
    const V3 X = eval3(S, M, u, v, 0);
    const V3 n = unit3(eval3(S, M, u, v, 9));
    const double ct = 0.02 + 0.98 * r3, sn = sqrt(1.0 - ct * ct), ph = 6.283185307179586 * r4;
    const V3 a = fabs(n.x) < 0.9 ? unit3(cross3(n, V3{ 1.0, 0.0, 0.0 })) : unit3(cross3(n, V3{ 0.0, 1.0, 0.0 }));
    const V3 b = cross3(n, a);
    const double cp = cos(ph), sp = sin(ph);
    const V3 d = { -(ct * n.x + sn * (cp * a.x + sp * b.x)), -(ct * n.y + sn * (cp * a.y + sp * b.y)),
                   -(ct * n.z + sn * (cp * a.z + sp * b.z)) };
    const double tau = frame[(size_t)p * kFrameDoubles + 12] * (0.25 + 3.75 * r5);
    Pair pr;
    pr.patch = (int32_t)p;
    pr.d = d;
    pr.o = { X.x - tau * d.x, X.y - tau * d.y, X.z - tau * d.z };
    return pr;
}

__device__ __forceinline__ Ray32 ray32_of(const V3& o, const V3& d) {
    return { (float)o.x, (float)o.y, (float)o.z, 1.0f / (float)d.x, 1.0f / (float)d.y, 1.0f / (float)d.z };
}

// ---------------------------------------------------------------------------
// Per-patch preparation
// ---------------------------------------------------------------------------
// Frame (centre X(0,0), unit normal from the stored normal field, e1 = X_u
// made orthogonal to it, e2 = n x e1, scale = largest distance of the 15
// feature points from the centre) and the 45 patch features.
__global__ void k_mlp_patch_prep(DevScene S, int64_t n, double* frame, float* pfeat) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const PatchMeta M = S.meta[k];
    const V3 c = eval3(S, M, 0.0, 0.0, 0);
    const V3 xu = eval3(S, M, 0.0, 0.0, 3);
    const V3 xv = eval3(S, M, 0.0, 0.0, 6);
    const V3 nu = eval3(S, M, 0.0, 0.0, 9);
    const double ln = sqrt(dot3(nu, nu));
    const V3 nn = { nu.x / ln, nu.y / ln, nu.z / ln };
    const double kk = dot3(xu, nn);
    const V3 w = { xu.x - kk * nn.x, xu.y - kk * nn.y, xu.z - kk * nn.z };
    const double lw = sqrt(dot3(w, w));
    const V3 e1 = { w.x / lw, w.y / lw, w.z / lw };
    const V3 e2 = cross3(nn, e1);
    const double fu[5] = { -1.0, -sqrt(3.0 / 7.0), 0.0, sqrt(3.0 / 7.0), 1.0 };
    const double fv[3] = { -1.0, 0.0, 1.0 };
    V3 P[15];
    double s = 0.0;
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 3; ++b) {
            const V3 x = eval3(S, M, fu[a], fv[b], 0);
            P[a * 3 + b] = x;
            const V3 dq = sub3(x, c);
            const double l = sqrt(dot3(dq, dq));
            s = (l > s) ? l : s;
        }
    double* F = frame + (size_t)k * kFrameDoubles;
    F[0] = c.x; F[1] = c.y; F[2] = c.z; F[3] = e1.x; F[4] = e1.y; F[5] = e1.z;
    F[6] = e2.x; F[7] = e2.y; F[8] = e2.z; F[9] = nn.x; F[10] = nn.y; F[11] = nn.z;
    F[12] = s; F[13] = 0.0; F[14] = 0.0; F[15] = 0.0;
    // rows of the pseudo-inverse of [X_u X_v]: tangent-plane coordinates u = a_u . (P - c), v = a_v . (P - c)
    const double guu = dot3(xu, xu), guv = dot3(xu, xv), gvv = dot3(xv, xv);
    const double det = guu * gvv - guv * guv;
    const double i00 = gvv / det, i01 = -guv / det, i11 = guu / det;
    F[16] = i00 * xu.x + i01 * xv.x; F[17] = i00 * xu.y + i01 * xv.y; F[18] = i00 * xu.z + i01 * xv.z;
    F[19] = i01 * xu.x + i11 * xv.x; F[20] = i01 * xu.y + i11 * xv.y; F[21] = i01 * xu.z + i11 * xv.z;
    F[22] = 0.0; F[23] = 0.0;
    float* f = pfeat + (size_t)k * kPatchFeatPad;
    for (int q = 0; q < 15; ++q) {
        const V3 dq = sub3(P[q], c);
        f[3 * q] = (float)(dot3(dq, e1) / s);
        f[3 * q + 1] = (float)(dot3(dq, e2) / s);
        f[3 * q + 2] = (float)(dot3(dq, nn) / s);
    }
    f[45] = 0.0f; f[46] = 0.0f; f[47] = 0.0f;
}

// The patch half of layer 1: h1pre[k][o] = b1[o] + sum_j bf16(W1[o][6+j]) bf16(f_k[j]) (fp32).
__global__ void k_mlp_h1pre(const float* pfeat, const float* P, int64_t n, float* h1pre) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * kH) return;
    const int64_t k = g / kH;
    const int o = (int)(g % kH);
    const float* f = pfeat + (size_t)k * kPatchFeatPad;
    const float* w = P + kOffW1 + o * kH + kRay;
    float acc = P[kOffB1 + o];
    for (int j = 0; j < kPatchF; ++j) acc += bf2f(f2bf(w[j])) * bf2f(f2bf(f[j]));
    h1pre[g] = acc;
}

// ---------------------------------------------------------------------------
// Pairs, features, A/B
// ---------------------------------------------------------------------------
// Features of a pair into a row of 64 bf16 (and optionally fp32 x[51]).
__device__ RayFeat write_features(const double* frame, const float* pfeat, int32_t p, const V3& o, const V3& d,
                                  uint16_t* row, float* xf) {
    const RayFeat rf = ray_features(frame + (size_t)p * kFrameDoubles, o, d);
    const float* r = rf.r;
    const float* f = pfeat + (size_t)p * kPatchFeatPad;
    if (row) {
        for (int e = 0; e < kRay; ++e) row[e] = (uint16_t)f2bf(r[e]);
        for (int e = 0; e < kPatchF; ++e) row[kRay + e] = (uint16_t)f2bf(f[e]);
        for (int e = kIn; e < 64; ++e) row[e] = 0;
    }
    if (xf) {
        for (int e = 0; e < kRay; ++e) xf[e] = r[e];
        for (int e = 0; e < kPatchF; ++e) xf[kRay + e] = f[e];
    }
    return rf;
}

// Pairs first .. first+n-1: the method's converged root as the target.
__global__ void k_mlp_gen(DevScene S, const double* frame, const float* pfeat, int64_t n_patch, uint64_t seed,
                          int64_t first, int64_t n, uint16_t* rows, float* target, int32_t* valid, int32_t* patch_out,
                          double* o_out, double* d_out, float* x_out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const Pair pr = draw_pair(S, frame, n_patch, seed, first + g);
    Counters c = {};
    const Seed sd = stage1_seed<2>(S, pr.patch, pr.o, pr.d, ray32_of(pr.o, pr.d), true, S.scenes[0].cull_widen, c);
    PHit h{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    if (sd.have) h = newton_any(S, pr.patch, pr.o, pr.d, sd.u, sd.v, sd.t, BBRT_NEWTON_CONVERGED, c);
    const bool ok = h.status == 1;
    const RayFeat rf = write_features(frame, pfeat, pr.patch, pr.o, pr.d, rows ? rows + (size_t)g * 64 : nullptr,
                                      x_out ? x_out + (size_t)g * kIn : nullptr);
    target[3 * g] = ok ? (float)(h.u - rf.ua) : 0.0f;
    target[3 * g + 1] = ok ? (float)(h.v - rf.va) : 0.0f;
    target[3 * g + 2] = ok ? (float)((h.t - rf.tc) / rf.s) : 0.0f;
    valid[g] = ok ? 1 : 0;
    if (patch_out) patch_out[g] = pr.patch;
    if (o_out) { o_out[3 * g] = pr.o.x; o_out[3 * g + 1] = pr.o.y; o_out[3 * g + 2] = pr.o.z; }
    if (d_out) { d_out[3 * g] = pr.d.x; d_out[3 * g + 1] = pr.d.y; d_out[3 * g + 2] = pr.d.z; }
}

__global__ void k_mlp_features(const double* frame, const float* pfeat, const int32_t* patch, const double* o,
                               const double* d, int64_t n, float* x, double* tc_out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const RayFeat rf = write_features(frame, pfeat, patch[g], V3{ o[3 * g], o[3 * g + 1], o[3 * g + 2] },
                                      V3{ d[3 * g], d[3 * g + 1], d[3 * g + 2] }, nullptr, x + (size_t)g * kIn);
    tc_out[3 * g] = rf.tc;
    tc_out[3 * g + 1] = rf.ua;
    tc_out[3 * g + 2] = rf.va;
}

// Seed A/B: per pair, converged Newton from the triangle seed and from the learned seed.
__global__ void __launch_bounds__(256) k_mlp_ab(DevScene S, MlpDev M, int64_t n_patch, uint64_t seed, int64_t first,
                                                int64_t n, int8_t* status, uint8_t* iters, uint8_t* accepted,
                                                float* seed_err, double* point_diff) {
    __shared__ MlpLds lds;
    mlp_stage_lds(M, lds, threadIdx.x, blockDim.x);
    __syncthreads();
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = g < n;
    Counters c = {};
    Pair pr{ 0, { 0, 0, 0 }, { 1, 0, 0 } };
    Seed sd{ 0.0, 0.0, 0.0, false };
    PHit ht{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    if (in) {
        pr = draw_pair(S, M.frame, n_patch, seed, first + g);
        sd = stage1_seed<2>(S, pr.patch, pr.o, pr.d, ray32_of(pr.o, pr.d), true, S.scenes[0].cull_widen, c);
        if (sd.have) ht = newton_any(S, pr.patch, pr.o, pr.d, sd.u, sd.v, sd.t, BBRT_NEWTON_CONVERGED, c);
    }
    double ul, vl, tl;
    mlp_seed_wave(M, lds, in && sd.have, pr.patch, pr.o, pr.d, ul, vl, tl);
    if (!in) return;
    PHit hl{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    if (sd.have) hl = newton_any(S, pr.patch, pr.o, pr.d, ul, vl, tl, BBRT_NEWTON_CONVERGED, c);
    status[g] = (int8_t)ht.status;
    status[n + g] = (int8_t)hl.status;
    iters[g] = (uint8_t)ht.iters;
    iters[n + g] = (uint8_t)hl.iters;
    const bool acc = hl.status == 1 && hl.polish_ok && fabs(hl.u - sd.u) <= kSeedWindow && fabs(hl.v - sd.v) <= kSeedWindow;
    accepted[g] = acc ? 1 : 0;
    const bool ref = ht.status == 1;
    seed_err[g] = ref ? (float)fmax(fabs(sd.u - ht.u), fabs(sd.v - ht.v)) : -1.0f;
    seed_err[n + g] = ref ? (float)fmax(fabs(ul - ht.u), fabs(vl - ht.v)) : -1.0f;
    const V3 dp = sub3(hl.p, ht.p);
    point_diff[g] = (ref && hl.status == 1) ? sqrt(dot3(dp, dp)) : -1.0;
}

// ---------------------------------------------------------------------------
// Training kernels
// ---------------------------------------------------------------------------
// fp32 master -> bf16 fragments
__global__ void k_mlp_pack(const float* __restrict__ P, u32x2_t* __restrict__ F) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= kNFrags * 64) return;
    const int f = g >> 6, l = g & 63, q = l >> 4, j = l & 15;
    float v[4];
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * q + i;
        float x;
        if (f < kFragFwd4) {            // W1..W3 forward: W[ot*16 + j][kt*16 + k]
            const int L = f / 16, ot = (f % 16) / 4, kt = f % 4;
            const int off = L == 0 ? kOffW1 : (L == 1 ? kOffW2 : kOffW3);
            x = P[off + (ot * 16 + j) * kH + kt * 16 + k];
        } else if (f < kFragBwd2) {     // W4 forward: W4[j][kt*16 + k]
            const int kt = f - kFragFwd4;
            x = P[kOffW4 + j * kH + kt * 16 + k];
        } else if (f < kFragBwd4) {     // W2^T, W3^T: W[ot*16 + k][it*16 + j]
            const int L = (f - kFragBwd2) / 16, it = ((f - kFragBwd2) % 16) / 4, ot = (f - kFragBwd2) % 4;
            const int off = L == 0 ? kOffW2 : kOffW3;
            x = P[off + (ot * 16 + k) * kH + it * 16 + j];
        } else if (f < kFragRay1) {     // W4^T: W4[k][it*16 + j]
            const int it = f - kFragBwd4;
            x = P[kOffW4 + k * kH + it * 16 + j];
        } else {                        // W1 ray columns: W1[ot*16 + j][k], k < 6
            const int ot = f - kFragRay1;
            x = k < kRay ? P[kOffW1 + (ot * 16 + j) * kH + k] : 0.0f;
        }
        v[i] = x;
    }
    F[g] = pack4(v[0], v[1], v[2], v[3]);
}

// fp32 features [n][51] -> bf16 rows [n][64]
__global__ void k_mlp_pack_x(const float* __restrict__ x, int64_t n, u32x2_t* __restrict__ X) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * kFeatRow) return;
    const int64_t s = g / kFeatRow;
    const int e = (int)(g % kFeatRow) * 4;
    float v[4];
    for (int i = 0; i < 4; ++i) v[i] = (e + i < kIn) ? x[s * kIn + e + i] : 0.0f;
    X[g] = pack4(v[0], v[1], v[2], v[3]);
}

constexpr int kLdsPitch = 68;   // feature-major rows of 64 samples, padded

struct TrainLds {
    uint16_t act[4][64][kLdsPitch];   // x, h1, h2, h3: [feature][sample]
    uint16_t dz[64][kLdsPitch];       // dZ of the layer being back-propagated: [feature][sample]
    f32x4_t gp[3][16][64];            // GELU'(z1..z3) in D order
};

// Stores B fragment (feature tile ft, sample tile st) feature-major.
__device__ __forceinline__ void store_fm(uint16_t (*A)[kLdsPitch], int ft, int st, u32x2_t v, int lane) {
    const int q = lane >> 4, j = lane & 15;
    const int col = st * 16 + j, row = ft * 16 + 4 * q;
    A[row][col] = (uint16_t)(v.x & 0xFFFFu);
    A[row + 1][col] = (uint16_t)(v.x >> 16);
    A[row + 2][col] = (uint16_t)(v.y & 0xFFFFu);
    A[row + 3][col] = (uint16_t)(v.y >> 16);
}

// Operand for a product contracting over samples: lane l reads row r0 + l%16,
// samples sk*16 + 4(l/16) .. +3.
__device__ __forceinline__ u32x2_t load_fm(const uint16_t (*A)[kLdsPitch], int r0, int sk, int lane) {
    const int q = lane >> 4, j = lane & 15;
    return *reinterpret_cast<const u32x2_t*>(&A[r0 + j][sk * 16 + 4 * q]);
}

// dW[out][in] = sum_s dZ[s][out] act[s][in] over the tile's 64 samples.
__device__ __forceinline__ void weight_grad(const uint16_t (*dz)[kLdsPitch], const uint16_t (*act)[kLdsPitch], int n_ot,
                                            float* part, int lane) {
    const int q = lane >> 4, j = lane & 15;
    for (int ot = 0; ot < n_ot; ++ot)
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            f32x4_t acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int sp = 0; sp < 2; ++sp)
                acc = mfma32(load_fm(dz, ot * 16, 2 * sp, lane), load_fm(dz, ot * 16, 2 * sp + 1, lane),
                             load_fm(act, it * 16, 2 * sp, lane), load_fm(act, it * 16, 2 * sp + 1, lane), acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) part[(ot * 16 + 4 * q + i) * kH + it * 16 + j] = acc[i];
        }
}

__device__ __forceinline__ void bias_grad(const uint16_t (*dz)[kLdsPitch], int n_out, float* part, int lane) {
    if (lane < n_out) {
        float s = 0.0f;
        for (int k = 0; k < 64; ++k) s += bf2f(dz[lane][k]);
        part[lane] = s;
    }
}

// dZ (B fragments, tile it x st) = (W^T dZ_next) * GELU'
template <int OT>
__device__ __forceinline__ void back_layer(const u32x2_t* Wb, const u32x2_t (&dzn)[OT][4], const f32x4_t* gp,
                                           u32x2_t (&dz)[4][4], int lane) {
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            f32x4_t acc = { 0.0f, 0.0f, 0.0f, 0.0f };
            if constexpr (OT % 2 == 0) {
#pragma unroll
                for (int op = 0; op < OT / 2; ++op)
                    acc = mfma32(Wb[(it * OT + 2 * op) * 64 + lane], Wb[(it * OT + 2 * op + 1) * 64 + lane],
                                 dzn[2 * op][st], dzn[2 * op + 1][st], acc);
            } else {
#pragma unroll
                for (int ot = 0; ot < OT; ++ot) acc = mfma16(Wb[(it * OT + ot) * 64 + lane], dzn[ot][st], acc);
            }
            const f32x4_t g = gp[(it * 4 + st) * 64 + lane];
            dz[it][st] = pack4(acc[0] * g[0], acc[1] * g[1], acc[2] * g[2], acc[3] * g[3]);
        }
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void load_tile(const u32x2_t* X, int64_t base, int64_t end, u32x2_t (&x)[4][4], int lane) {
    const int q = lane >> 4, j = lane & 15;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const int64_t smp = base + st * 16 + j;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            u32x2_t z;
            z.x = 0u; z.y = 0u;
            x[kt][st] = smp < end ? X[smp * kFeatRow + kt * 4 + q] : z;
        }
    }
}

// One 64-sample tile: forward, loss, backward; the tile's gradient to part.
__global__ void __launch_bounds__(64) k_mlp_train(const u32x2_t* __restrict__ X, const float* __restrict__ T,
                                                  int64_t first, int64_t n, float scale, const u32x2_t* __restrict__ F,
                                                  const float* __restrict__ P, float* __restrict__ partial,
                                                  float* __restrict__ loss_part) {
    __shared__ TrainLds L;
    const int lane = threadIdx.x, q = lane >> 4, j = lane & 15;
    const int64_t base = first + (int64_t)blockIdx.x * 64, end = first + n;
    float* part = partial + (size_t)blockIdx.x * kNParams;
    u32x2_t x[4][4];
    load_tile(X, base, end, x, lane);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int st = 0; st < 4; ++st) store_fm(L.act[0], kt, st, x[kt][st], lane);
    u32x2_t h1[4][4], h2[4][4], h3[4][4];
    hidden_layer<true>(F + kFragFwd1 * 64, P + kOffB1, x, h1, &L.gp[0][0][0], lane);
#pragma unroll
    for (int ft = 0; ft < 4; ++ft)
#pragma unroll
        for (int st = 0; st < 4; ++st) store_fm(L.act[1], ft, st, h1[ft][st], lane);
    hidden_layer<true>(F + kFragFwd2 * 64, P + kOffB2, h1, h2, &L.gp[1][0][0], lane);
#pragma unroll
    for (int ft = 0; ft < 4; ++ft)
#pragma unroll
        for (int st = 0; st < 4; ++st) store_fm(L.act[2], ft, st, h2[ft][st], lane);
    hidden_layer<true>(F + kFragFwd3 * 64, P + kOffB3, h2, h3, &L.gp[2][0][0], lane);
#pragma unroll
    for (int ft = 0; ft < 4; ++ft)
#pragma unroll
        for (int st = 0; st < 4; ++st) store_fm(L.act[3], ft, st, h3[ft][st], lane);
    f32x4_t y[4];
    out_layer(F + kFragFwd4 * 64, P + kOffB4, h3, y, lane);

    // loss and dY (lanes 0..15 hold outputs 0..3 of their samples)
    u32x2_t dy[1][4];
    float lsum = 0.0f;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const int64_t smp = base + st * 16 + j;
        float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
        if (q == 0 && smp < end) {
            const float r0 = y[st][0] - T[3 * smp], r1 = y[st][1] - T[3 * smp + 1], r2 = y[st][2] - T[3 * smp + 2];
            lsum += r0 * r0 + r1 * r1 + r2 * r2;
            d0 = r0 * scale; d1 = r1 * scale; d2 = r2 * scale;
        }
        dy[0][st] = pack4(d0, d1, d2, 0.0f);
        store_fm(L.dz, 0, st, dy[0][st], lane);
    }
    __syncthreads();
    weight_grad(L.dz, L.act[3], 1, part + kOffW4, lane);
    bias_grad(L.dz, kOutPad, part + kOffB4, lane);

    u32x2_t dz3[4][4], dz2[4][4], dz1[4][4];
    back_layer<1>(F + kFragBwd4 * 64, dy, &L.gp[2][0][0], dz3, lane);
    __syncthreads();
#pragma unroll
    for (int ft = 0; ft < 4; ++ft)
#pragma unroll
        for (int st = 0; st < 4; ++st) store_fm(L.dz, ft, st, dz3[ft][st], lane);
    __syncthreads();
    weight_grad(L.dz, L.act[2], 4, part + kOffW3, lane);
    bias_grad(L.dz, kH, part + kOffB3, lane);

    back_layer<4>(F + kFragBwd3 * 64, dz3, &L.gp[1][0][0], dz2, lane);
    __syncthreads();
#pragma unroll
    for (int ft = 0; ft < 4; ++ft)
#pragma unroll
        for (int st = 0; st < 4; ++st) store_fm(L.dz, ft, st, dz2[ft][st], lane);
    __syncthreads();
    weight_grad(L.dz, L.act[1], 4, part + kOffW2, lane);
    bias_grad(L.dz, kH, part + kOffB2, lane);

    back_layer<4>(F + kFragBwd2 * 64, dz2, &L.gp[0][0][0], dz1, lane);
    __syncthreads();
#pragma unroll
    for (int ft = 0; ft < 4; ++ft)
#pragma unroll
        for (int st = 0; st < 4; ++st) store_fm(L.dz, ft, st, dz1[ft][st], lane);
    __syncthreads();
    weight_grad(L.dz, L.act[0], 4, part + kOffW1, lane);
    bias_grad(L.dz, kH, part + kOffB1, lane);

    lsum = wave_sum_f(lsum);
    if (lane == 0) loss_part[blockIdx.x] = lsum;
}

// Forward only (outputs and/or summed squared error per tile).
__global__ void __launch_bounds__(64) k_mlp_forward(const u32x2_t* __restrict__ X, const float* __restrict__ T,
                                                    int64_t first, int64_t n, const u32x2_t* __restrict__ F,
                                                    const float* __restrict__ P, float* __restrict__ Y,
                                                    float* __restrict__ loss_part) {
    const int lane = threadIdx.x, q = lane >> 4, j = lane & 15;
    const int64_t base = first + (int64_t)blockIdx.x * 64, end = first + n;
    u32x2_t x[4][4], h1[4][4], h2[4][4], h3[4][4];
    load_tile(X, base, end, x, lane);
    hidden_layer<false>(F + kFragFwd1 * 64, P + kOffB1, x, h1, nullptr, lane);
    hidden_layer<false>(F + kFragFwd2 * 64, P + kOffB2, h1, h2, nullptr, lane);
    hidden_layer<false>(F + kFragFwd3 * 64, P + kOffB3, h2, h3, nullptr, lane);
    f32x4_t y[4];
    out_layer(F + kFragFwd4 * 64, P + kOffB4, h3, y, lane);
    float lsum = 0.0f;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const int64_t smp = base + st * 16 + j;
        if (q == 0 && smp < end) {
            if (Y) { Y[3 * (smp - first)] = y[st][0]; Y[3 * (smp - first) + 1] = y[st][1]; Y[3 * (smp - first) + 2] = y[st][2]; }
            if (T) {
                const float r0 = y[st][0] - T[3 * smp], r1 = y[st][1] - T[3 * smp + 1], r2 = y[st][2] - T[3 * smp + 2];
                lsum += r0 * r0 + r1 * r1 + r2 * r2;
            }
        }
    }
    if (loss_part) {
        lsum = wave_sum_f(lsum);
        if (lane == 0) loss_part[blockIdx.x] = lsum;
    }
}

// Sum of the tiles' gradients (fixed order) and the batch loss.
__global__ void k_mlp_reduce(const float* __restrict__ partial, int n_tiles, float* __restrict__ grad,
                             const float* __restrict__ loss_part, float inv_n, float* __restrict__ loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNParams) {
        float s = 0.0f;
        for (int t = 0; t < n_tiles; ++t) s += partial[(size_t)t * kNParams + p];
        grad[p] = s;
    }
    if (p == 0 && loss_out) {
        float s = 0.0f;
        for (int t = 0; t < n_tiles; ++t) s += loss_part[t];
        *loss_out = s * inv_n;
    }
}

// One weight's bf16 copy into every fragment slot that holds it (the inverse of
// k_mlp_pack's gather): forward fragments of W1..W4, the W1 ray fragments, and
// the transposed fragments of W2..W4 the backward pass reads.
__device__ __forceinline__ void scatter_weight(uint16_t* __restrict__ F16, int p, uint16_t v) {
    auto put = [&](int f, int lane, int slot) { F16[((size_t)f * 64 + lane) * 4 + slot] = v; };
    if (p < kOffB1) {
        const int r = (p - kOffW1) / kH, c = (p - kOffW1) % kH;
        put(kFragFwd1 + (r / 16) * 4 + c / 16, ((c % 16) / 4) * 16 + r % 16, c % 4);
        if (c < kRay) put(kFragRay1 + r / 16, (c / 4) * 16 + r % 16, c % 4);
    } else if ((p >= kOffW2 && p < kOffB2) || (p >= kOffW3 && p < kOffB3)) {
        const int L = p < kOffB2 ? 0 : 1;
        const int off = L == 0 ? kOffW2 : kOffW3;
        const int r = (p - off) / kH, c = (p - off) % kH;
        put((L == 0 ? kFragFwd2 : kFragFwd3) + (r / 16) * 4 + c / 16, ((c % 16) / 4) * 16 + r % 16, c % 4);
        put(kFragBwd2 + L * 16 + (c / 16) * 4 + r / 16, ((r % 16) / 4) * 16 + c % 16, r % 4);
    } else if (p >= kOffW4 && p < kOffB4) {
        const int r = (p - kOffW4) / kH, c = (p - kOffW4) % kH;
        put(kFragFwd4 + c / 16, ((c % 16) / 4) * 16 + r, c % 4);
        put(kFragBwd4 + c / 16, (r / 4) * 16 + c % 16, r % 4);
    }
}

// Adam on the fp32 master weights (the pads have zero gradient and stay zero)
// fused with their repacking.  A workgroup owns kAdamParams parameters and
// sums their tile gradients with kAdamGroups waves: wave w adds tiles w,
// w + kAdamGroups, ... in order, then wave 0 adds the waves' sums in order, so
// the step is deterministic.  Wave 0 then updates each parameter and writes its
// bf16 value straight into the fragments (the slots no parameter owns -- the
// ray fragments' zero columns -- keep what k_mlp_pack first wrote).
constexpr int kAdamParams = 64, kAdamGroups = 8;

__global__ void __launch_bounds__(kAdamParams * kAdamGroups)
k_mlp_adam_pack(float* __restrict__ P, float* __restrict__ m1, float* __restrict__ m2, const float* __restrict__ partial,
                int n_tiles, float lr, float bc1, float bc2, const float* __restrict__ loss_part, float inv_n,
                float* __restrict__ loss_hist, uint16_t* __restrict__ F16) {
    __shared__ float sums[kAdamGroups][kAdamParams];
    const int pl = threadIdx.x % kAdamParams, grp = threadIdx.x / kAdamParams;
    const int p = blockIdx.x * kAdamParams + pl;
    float g = 0.0f;
    if (p < kNParams)
        for (int t = grp; t < n_tiles; t += kAdamGroups) g += partial[(size_t)t * kNParams + p];
    sums[grp][pl] = g;
    __syncthreads();
    if (grp != 0 || p >= kNParams) return;
    g = 0.0f;
#pragma unroll
    for (int w = 0; w < kAdamGroups; ++w) g += sums[w][pl];
    const float a = 0.9f * m1[p] + 0.1f * g;
    const float b = 0.999f * m2[p] + 0.001f * g * g;
    m1[p] = a;
    m2[p] = b;
    const float w = P[p] - lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
    P[p] = w;
    scatter_weight(F16, p, (uint16_t)f2bf(w));
    if (p == 0 && loss_hist) {
        float s = 0.0f;
        for (int t = 0; t < n_tiles; ++t) s += loss_part[t];
        *loss_hist = s * inv_n;
    }
}

// Deterministic compaction of the valid pairs (block counts, one-block scan, scatter).
__global__ void k_count_valid(const int32_t* valid, int64_t n, int32_t* blk) {
    __shared__ int32_t w[4];
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const unsigned long long m = __ballot(g < n && valid[g]);
    if (__lane_id() == 0) w[threadIdx.x >> 6] = (int32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

__global__ void k_scan_blocks(int32_t* blk, int64_t nb, int64_t* total) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024, lo = t * per, hi = std::min<int64_t>(nb, lo + per);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += blk[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int i = 0; i < 1024; ++i) { const int64_t v = part[i]; part[i] = run; run += v; }
        *total = run;
    }
    __syncthreads();
    int64_t run = part[t];
    for (int64_t i = lo; i < hi; ++i) { const int32_t v = blk[i]; blk[i] = (int32_t)run; run += v; }
}

__global__ void k_compact(const int32_t* valid, int64_t n, const int32_t* blk_off, const u32x2_t* rows_in,
                          const float* t_in, u32x2_t* rows_out, float* t_out) {
    __shared__ int32_t w[4];
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool v = g < n && valid[g];
    const unsigned long long m = __ballot(v);
    const int wid = threadIdx.x >> 6, lane = __lane_id();
    if (lane == 0) w[wid] = (int32_t)__popcll(m);
    __syncthreads();
    int off = blk_off[blockIdx.x];
    for (int k = 0; k < wid; ++k) off += w[k];
    if (v) {
        const int64_t dst = off + (int64_t)__popcll(m & ((1ull << lane) - 1ull));
        for (int e = 0; e < kFeatRow; ++e) rows_out[dst * kFeatRow + e] = rows_in[g * kFeatRow + e];
        t_out[3 * dst] = t_in[3 * g]; t_out[3 * dst + 1] = t_in[3 * g + 1]; t_out[3 * dst + 2] = t_in[3 * g + 2];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
uint64_t host_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <class T>
T* dmalloc(size_t count) {
    T* p = nullptr;
    check(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)), "hipMalloc");
    return p;
}

struct DevFree {
    std::vector<void*> ps;
    ~DevFree() { for (void* p : ps) (void)hipFree(p); }
    template <class T>
    T* get(size_t count) { T* p = dmalloc<T>(count); ps.push_back(p); return p; }
};

unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct bbrt_mlp_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<float> host;   // master parameters, kept in sync with P
    float* P = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    u32x2_t* F = nullptr;
    int64_t steps = 0;
    ~bbrt_mlp_s() {
        for (void* p : { (void*)P, (void*)m1, (void*)m2, (void*)F }) if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void mlp_upload(bbrt_mlp m) {
    check(hipMemcpyAsync(m->P, m->host.data(), kNParams * sizeof(float), hipMemcpyHostToDevice, m->stream), "params H2D");
    hipLaunchKernelGGL(k_mlp_pack, dim3(blocks_for(kNFrags * 64, 256)), dim3(256), 0, m->stream, m->P, m->F);
    check(hipGetLastError(), "k_mlp_pack");
    check(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
}

void mlp_download(bbrt_mlp m) {
    check(hipMemcpyAsync(m->host.data(), m->P, kNParams * sizeof(float), hipMemcpyDeviceToHost, m->stream), "params D2H");
    check(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
}

bool is_pad(int p) {
    if (p >= kOffW1 && p < kOffB1) return (p - kOffW1) % kH >= kIn;
    if (p >= kOffW4 && p < kOffB4) return (p - kOffW4) / kH >= kOut;
    if (p >= kOffB4) return p - kOffB4 >= kOut;
    return false;
}

// Per-patch frames and features of a scene (weight-independent; computed once).
void ensure_frames(bbrt_scene s) {
    if (s->mlp_frame) return;
    check(hipSetDevice(s->device), "hipSetDevice");
    const int64_t n = s->n_patches;
    s->mlp_frame = dmalloc<double>((size_t)n * kFrameDoubles);
    s->allocs.push_back(s->mlp_frame);
    s->mlp_pfeat = dmalloc<float>((size_t)n * kPatchFeatPad);
    s->allocs.push_back(s->mlp_pfeat);
    s->bytes += n * (kFrameDoubles * 8 + kPatchFeatPad * 4);
    hipLaunchKernelGGL(k_mlp_patch_prep, dim3(blocks_for(n, 64)), dim3(64), 0, s->stream, s->dev, n, s->mlp_frame,
                       s->mlp_pfeat);
    check(hipGetLastError(), "k_mlp_patch_prep");
    check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}

// Uploads host x [n][51] as bf16 rows.
u32x2_t* upload_rows(DevFree& mem, const float* x, int64_t n, hipStream_t st) {
    float* dx = mem.get<float>((size_t)n * kIn);
    check(hipMemcpyAsync(dx, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, st), "x H2D");
    u32x2_t* X = mem.get<u32x2_t>((size_t)n * kFeatRow);
    hipLaunchKernelGGL(k_mlp_pack_x, dim3(blocks_for(n * kFeatRow, 256)), dim3(256), 0, st, dx, n, X);
    check(hipGetLastError(), "k_mlp_pack_x");
    return X;
}

// Summed squared error of the network over rows [first, first+n) (double, host).
double split_loss(bbrt_mlp m, const u32x2_t* X, const float* T, int64_t first, int64_t n) {
    if (n <= 0) return 0.0;
    DevFree mem;
    const int64_t tiles = (n + 63) / 64;
    float* lp = mem.get<float>((size_t)tiles);
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)tiles), dim3(64), 0, m->stream, X, T, first, n, m->F, m->P,
                       nullptr, lp);
    check(hipGetLastError(), "k_mlp_forward");
    std::vector<float> h((size_t)tiles);
    check(hipMemcpyAsync(h.data(), lp, tiles * sizeof(float), hipMemcpyDeviceToHost, m->stream), "loss D2H");
    check(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    double s = 0.0;
    for (float v : h) s += v;
    return s / (double)n;
}

}  // namespace

extern "C" {

int bbrt_mlp_create(uint64_t seed, bbrt_mlp* out) {
    if (!out) { g_last_error = "bbrt_mlp_create: out is NULL"; return -1; }
    *out = nullptr;
    bbrt_mlp m = nullptr;
    try {
        m = new bbrt_mlp_s();
        check(hipGetDevice(&m->device), "hipGetDevice");
        check(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "hipStreamCreate");
        m->host.assign(kNParams, 0.0f);
        // Glorot-uniform weights from SplitMix64(seed): W1 (64 x 51), W2, W3 (64 x 64), W4 (3 x 64), row-major
        uint64_t k = 0;
        auto draw = [&](double a) {
            ++k;
            const double u = (double)(host_mix(seed + 0x9E3779B97F4A7C15ull * k) >> 11) * (1.0 / 9007199254740992.0);
            return (float)(-a + (a - -a) * u);
        };
        const struct { int off, rows, cols, fan_in, fan_out; } Ls[4] = {
            { kOffW1, kH, kIn, kIn, kH }, { kOffW2, kH, kH, kH, kH }, { kOffW3, kH, kH, kH, kH }, { kOffW4, kOut, kH, kH, kOut } };
        for (const auto& L : Ls) {
            const double a = std::sqrt(6.0 / (double)(L.fan_in + L.fan_out));
            for (int r = 0; r < L.rows; ++r)
                for (int c = 0; c < L.cols; ++c) m->host[L.off + r * kH + c] = draw(a);
        }
        m->P = dmalloc<float>(kNParams);
        m->m1 = dmalloc<float>(kNParams);
        m->m2 = dmalloc<float>(kNParams);
        m->F = dmalloc<u32x2_t>(kNFrags * 64);
        check(hipMemsetAsync(m->m1, 0, kNParams * sizeof(float), m->stream), "hipMemset");
        check(hipMemsetAsync(m->m2, 0, kNParams * sizeof(float), m->stream), "hipMemset");
        mlp_upload(m);
        *out = m;
        return 0;
    } catch (const std::exception& e) {
        delete m;
        return fail(e);
    }
}

void bbrt_mlp_destroy(bbrt_mlp m) { delete m; }

int bbrt_mlp_get_params(bbrt_mlp m, float* params) {
    if (!m || !params) { g_last_error = "bbrt_mlp_get_params: NULL argument"; return -1; }
    std::memcpy(params, m->host.data(), kNParams * sizeof(float));
    return 0;
}

int bbrt_mlp_set_params(bbrt_mlp m, const float* params) {
    if (!m || !params) { g_last_error = "bbrt_mlp_set_params: NULL argument"; return -1; }
    for (int p = 0; p < kNParams; ++p)
        if (is_pad(p) && params[p] != 0.0f) { g_last_error = "bbrt_mlp_set_params: padding parameters must be zero"; return -1; }
    try {
        check(hipSetDevice(m->device), "hipSetDevice");
        std::memcpy(m->host.data(), params, kNParams * sizeof(float));
        check(hipMemsetAsync(m->m1, 0, kNParams * sizeof(float), m->stream), "hipMemset");
        check(hipMemsetAsync(m->m2, 0, kNParams * sizeof(float), m->stream), "hipMemset");
        m->steps = 0;
        mlp_upload(m);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_mlp_forward(bbrt_mlp m, const float* x, int64_t n, float* y) {
    if (!m || (n > 0 && (!x || !y))) { g_last_error = "bbrt_mlp_forward: NULL argument"; return -1; }
    if (n < 0 || n > (int64_t)1 << 26) { g_last_error = "bbrt_mlp_forward: n out of range"; return -1; }
    if (n == 0) return 0;
    try {
        check(hipSetDevice(m->device), "hipSetDevice");
        DevFree mem;
        u32x2_t* X = upload_rows(mem, x, n, m->stream);
        float* Y = mem.get<float>((size_t)n * 3);
        hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(n, 64)), dim3(64), 0, m->stream, X, nullptr, 0, n, m->F, m->P,
                           Y, nullptr);
        check(hipGetLastError(), "k_mlp_forward");
        check(hipMemcpyAsync(y, Y, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, m->stream), "y D2H");
        check(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_mlp_loss_grad(bbrt_mlp m, const float* x, const float* target, int64_t n, float* loss, float* grad) {
    if (!m || !x || !target || !grad) { g_last_error = "bbrt_mlp_loss_grad: NULL argument"; return -1; }
    if (n <= 0 || n > (int64_t)1 << 20) { g_last_error = "bbrt_mlp_loss_grad: n must be in 1 .. 2^20"; return -1; }
    try {
        check(hipSetDevice(m->device), "hipSetDevice");
        DevFree mem;
        u32x2_t* X = upload_rows(mem, x, n, m->stream);
        float* T = mem.get<float>((size_t)n * 3);
        check(hipMemcpyAsync(T, target, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, m->stream), "t H2D");
        const int64_t tiles = (n + 63) / 64;
        float* part = mem.get<float>((size_t)tiles * kNParams);
        float* lp = mem.get<float>((size_t)tiles);
        float* g = mem.get<float>(kNParams);
        float* l = mem.get<float>(1);
        hipLaunchKernelGGL(k_mlp_train, dim3((unsigned)tiles), dim3(64), 0, m->stream, X, T, 0, n, 2.0f / (float)n, m->F,
                           m->P, part, lp);
        check(hipGetLastError(), "k_mlp_train");
        hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks_for(kNParams, 256)), dim3(256), 0, m->stream, part, (int)tiles, g, lp,
                           1.0f / (float)n, l);
        check(hipGetLastError(), "k_mlp_reduce");
        check(hipMemcpyAsync(grad, g, kNParams * sizeof(float), hipMemcpyDeviceToHost, m->stream), "grad D2H");
        float hl = 0.0f;
        check(hipMemcpyAsync(&hl, l, sizeof(float), hipMemcpyDeviceToHost, m->stream), "loss D2H");
        check(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
        if (loss) *loss = hl;
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_mlp_train(bbrt_mlp m, bbrt_scene s, const bbrt_train_config* cfg, float* loss_curve, bbrt_train_report* rep) {
    if (!m || !s || !cfg) { g_last_error = "bbrt_mlp_train: NULL argument"; return -1; }
    if (cfg->n_samples <= 0 || cfg->n_samples > ((int64_t)1 << 31) - 1024 || cfg->batch <= 0 || cfg->batch % 64 ||
        cfg->batch > (1 << 20) || cfg->epochs <= 0 || !(cfg->lr > 0.0f) || cfg->lr_final < 0.0f ||
        cfg->val_frac < 0.0f || cfg->test_frac < 0.0f || cfg->val_frac + cfg->test_frac >= 1.0f) {
        g_last_error = "bbrt_mlp_train: invalid configuration";
        return -1;
    }
    if (s->n_scenes != 1) { g_last_error = "bbrt_mlp_train: draws pairs from a single scene"; return -1; }
    try {
        check(hipSetDevice(s->device), "hipSetDevice");
        if (m->device != s->device) throw std::runtime_error("bbrt_mlp_train: network and scene on different devices");
        ensure_frames(s);
        hipStream_t st = m->stream;
        check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        DevFree mem;
        const int64_t n = cfg->n_samples;
        const auto t0 = std::chrono::steady_clock::now();
        u32x2_t* rows;
        float* tgt;
        int64_t n_pairs = 0;
        {
            DevFree raw;
            u32x2_t* rows_raw = raw.get<u32x2_t>((size_t)n * kFeatRow);
            float* t_raw = raw.get<float>((size_t)n * 3);
            int32_t* valid = raw.get<int32_t>((size_t)n);
            hipLaunchKernelGGL(k_mlp_gen, dim3(blocks_for(n, 256)), dim3(256), 0, st, s->dev, s->mlp_frame, s->mlp_pfeat,
                               s->n_patches, cfg->seed, (int64_t)0, n, reinterpret_cast<uint16_t*>(rows_raw), t_raw,
                               valid, nullptr, nullptr, nullptr, nullptr);
            check(hipGetLastError(), "k_mlp_gen");
            const int64_t nb = (n + 255) / 256;
            int32_t* blk = raw.get<int32_t>((size_t)nb);
            int64_t* total = raw.get<int64_t>(1);
            hipLaunchKernelGGL(k_count_valid, dim3((unsigned)nb), dim3(256), 0, st, valid, n, blk);
            hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, blk, nb, total);
            check(hipMemcpyAsync(&n_pairs, total, sizeof(int64_t), hipMemcpyDeviceToHost, st), "count D2H");
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            rows = mem.get<u32x2_t>((size_t)std::max<int64_t>(1, n_pairs) * kFeatRow);
            tgt = mem.get<float>((size_t)std::max<int64_t>(1, n_pairs) * 3);
            hipLaunchKernelGGL(k_compact, dim3((unsigned)nb), dim3(256), 0, st, valid, n, blk, rows_raw, t_raw, rows, tgt);
            check(hipGetLastError(), "k_compact");
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
        }
        const auto t1 = std::chrono::steady_clock::now();
        const int64_t n_test = (int64_t)std::floor((double)n_pairs * cfg->test_frac);
        const int64_t n_val = (int64_t)std::floor((double)n_pairs * cfg->val_frac);
        const int64_t n_train = n_pairs - n_val - n_test;
        const int64_t B = cfg->batch, per_epoch = n_train / B;
        if (per_epoch <= 0) throw std::runtime_error("bbrt_mlp_train: fewer training pairs than one batch");
        const int64_t total_steps = per_epoch * cfg->epochs;
        const int tiles = (int)(B / 64);
        float* part = mem.get<float>((size_t)tiles * kNParams);
        float* lp = mem.get<float>((size_t)tiles);
        float* hist = mem.get<float>((size_t)total_steps);
        const double b1 = 0.9, b2 = 0.999;
        for (int64_t t = 0; t < total_steps; ++t) {
            const int64_t first = (t % per_epoch) * B;
            const double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(M_PI * (double)t / (double)total_steps));
            const int64_t step = m->steps + t + 1;
            const float bc1 = (float)(1.0 - std::pow(b1, (double)step)), bc2 = (float)(1.0 - std::pow(b2, (double)step));
            hipLaunchKernelGGL(k_mlp_train, dim3(tiles), dim3(64), 0, st, rows, tgt, first, B, 2.0f / (float)B, m->F, m->P,
                               part, lp);
            hipLaunchKernelGGL(k_mlp_adam_pack, dim3(blocks_for(kNParams, kAdamParams)), dim3(kAdamParams * kAdamGroups), 0,
                               st, m->P, m->m1, m->m2,
                               part, tiles, (float)lr, bc1, bc2, lp, 1.0f / (float)B, hist + t,
                               reinterpret_cast<uint16_t*>(m->F));
        }
        check(hipGetLastError(), "training launches");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        m->steps += total_steps;
        const auto t2 = std::chrono::steady_clock::now();
        std::vector<float> h((size_t)total_steps);
        check(hipMemcpyAsync(h.data(), hist, total_steps * sizeof(float), hipMemcpyDeviceToHost, st), "loss D2H");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (loss_curve) std::memcpy(loss_curve, h.data(), h.size() * sizeof(float));
        mlp_download(m);
        if (rep) {
            rep->n_pairs = n_pairs; rep->n_train = n_train; rep->n_val = n_val; rep->n_test = n_test;
            rep->steps = total_steps;
            rep->first_loss = h.front(); rep->final_loss = h.back();
            rep->val_loss = (float)split_loss(m, rows, tgt, n_train, n_val);
            rep->test_loss = (float)split_loss(m, rows, tgt, n_train + n_val, n_test);
            rep->gen_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            rep->train_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_mlp_sample_pairs(bbrt_scene s, uint64_t seed, int64_t first, int64_t n, int32_t* patch, double* origin,
                          double* dir, float* x, float* target, int8_t* valid) {
    if (!s || !patch || !origin || !dir || !x || !target || !valid) {
        g_last_error = "bbrt_mlp_sample_pairs: NULL argument";
        return -1;
    }
    if (n < 0 || n > (int64_t)1 << 24 || first < 0) { g_last_error = "bbrt_mlp_sample_pairs: n out of range"; return -1; }
    if (s->n_scenes != 1) { g_last_error = "bbrt_mlp_sample_pairs: draws pairs from a single scene"; return -1; }
    if (n == 0) return 0;
    try {
        ensure_frames(s);
        DevFree mem;
        float* t = mem.get<float>((size_t)n * 3);
        int32_t* v = mem.get<int32_t>((size_t)n);
        int32_t* pp = mem.get<int32_t>((size_t)n);
        double* oo = mem.get<double>((size_t)n * 3);
        double* dd = mem.get<double>((size_t)n * 3);
        float* xx = mem.get<float>((size_t)n * kIn);
        hipLaunchKernelGGL(k_mlp_gen, dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, s->dev, s->mlp_frame, s->mlp_pfeat,
                           s->n_patches, seed, first, n, nullptr, t, v, pp, oo, dd, xx);
        check(hipGetLastError(), "k_mlp_gen");
        std::vector<int32_t> hv((size_t)n);
        check(hipMemcpyAsync(patch, pp, n * sizeof(int32_t), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(origin, oo, n * 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(dir, dd, n * 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(x, xx, n * kIn * sizeof(float), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(target, t, n * 3 * sizeof(float), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(hv.data(), v, n * sizeof(int32_t), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        for (int64_t i = 0; i < n; ++i) valid[i] = (int8_t)hv[i];
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_mlp_features(bbrt_scene s, const int32_t* patch, const double* origin, const double* dir, int64_t n, float* x,
                      double* tc, double* frame) {
    if (!s || (n > 0 && (!patch || !origin || !dir || !x || !tc))) { g_last_error = "bbrt_mlp_features: NULL argument"; return -1; }
    if (n < 0 || n > (int64_t)1 << 24) { g_last_error = "bbrt_mlp_features: n out of range"; return -1; }
    for (int64_t i = 0; i < n; ++i)
        if (patch[i] < 0 || patch[i] >= s->n_patches) { g_last_error = "bbrt_mlp_features: patch index out of range"; return -1; }
    try {
        ensure_frames(s);
        if (n > 0) {
            DevFree mem;
            int32_t* pp = mem.get<int32_t>((size_t)n);
            double* oo = mem.get<double>((size_t)n * 3);
            double* dd = mem.get<double>((size_t)n * 3);
            float* xx = mem.get<float>((size_t)n * kIn);
            double* tt = mem.get<double>((size_t)n * 3);
            check(hipMemcpyAsync(pp, patch, n * sizeof(int32_t), hipMemcpyHostToDevice, s->stream), "H2D");
            check(hipMemcpyAsync(oo, origin, n * 3 * sizeof(double), hipMemcpyHostToDevice, s->stream), "H2D");
            check(hipMemcpyAsync(dd, dir, n * 3 * sizeof(double), hipMemcpyHostToDevice, s->stream), "H2D");
            hipLaunchKernelGGL(k_mlp_features, dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, s->mlp_frame,
                               s->mlp_pfeat, pp, oo, dd, n, xx, tt);
            check(hipGetLastError(), "k_mlp_features");
            check(hipMemcpyAsync(x, xx, n * kIn * sizeof(float), hipMemcpyDeviceToHost, s->stream), "D2H");
            check(hipMemcpyAsync(tc, tt, n * 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream), "D2H");
        }
        if (frame)
            check(hipMemcpyAsync(frame, s->mlp_frame, (size_t)s->n_patches * kFrameDoubles * sizeof(double),
                                 hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_scene_set_mlp(bbrt_scene s, bbrt_mlp m) {
    if (!s || !m) { g_last_error = "bbrt_scene_set_mlp: NULL argument"; return -1; }
    try {
        // No asynchronous trace may be enqueued meanwhile, and every trace already in flight on the device
        // (bbrt_trace_device_async / _noalloc, on any stream) ends before the network buffers are rewritten:
        // it reads them through the whole launch.
        std::lock_guard<std::mutex> lock(s->async_mu);
        check(hipSetDevice(s->device), "hipSetDevice");
        if (m->device != s->device) throw std::runtime_error("bbrt_scene_set_mlp: network and scene on different devices");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        ensure_frames(s);
        check(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
        if (!s->mlp_h1pre) {
            s->mlp_h1pre = dmalloc<float>((size_t)s->n_patches * kH);
            s->allocs.push_back(s->mlp_h1pre);
            s->mlp_wfrag = dmalloc<u32x2_t>(kNFrags * 64);
            s->allocs.push_back(s->mlp_wfrag);
            s->mlp_params = dmalloc<float>(kNParams);
            s->allocs.push_back(s->mlp_params);
            s->bytes += s->n_patches * kH * 4 + kNFrags * 64 * 8 + kNParams * 4;
        }
        check(hipMemcpyAsync(s->mlp_wfrag, m->F, kNFrags * 64 * sizeof(u32x2_t), hipMemcpyDeviceToDevice, s->stream), "D2D");
        check(hipMemcpyAsync(s->mlp_params, m->P, kNParams * sizeof(float), hipMemcpyDeviceToDevice, s->stream), "D2D");
        hipLaunchKernelGGL(k_mlp_h1pre, dim3(blocks_for(s->n_patches * kH, 256)), dim3(256), 0, s->stream, s->mlp_pfeat,
                           s->mlp_params, s->n_patches, s->mlp_h1pre);
        check(hipGetLastError(), "k_mlp_h1pre");
        check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        s->mlp_ready = true;
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_mlp_seed_ab(bbrt_scene s, uint64_t seed, int64_t first, int64_t n, int8_t* status, uint8_t* iters,
                     uint8_t* accepted, float* seed_err, double* point_diff) {
    if (!s || (n > 0 && (!status || !iters || !accepted || !seed_err || !point_diff))) {
        g_last_error = "bbrt_mlp_seed_ab: NULL argument";
        return -1;
    }
    if (!s->mlp_ready) { g_last_error = "bbrt_mlp_seed_ab: no network attached (bbrt_scene_set_mlp)"; return -1; }
    if (n < 0 || n > (int64_t)1 << 26 || first < 0) { g_last_error = "bbrt_mlp_seed_ab: n out of range"; return -1; }
    if (s->n_scenes != 1) { g_last_error = "bbrt_mlp_seed_ab: draws pairs from a single scene"; return -1; }
    if (n == 0) return 0;
    try {
        check(hipSetDevice(s->device), "hipSetDevice");
        DevFree mem;
        int8_t* st = mem.get<int8_t>((size_t)n * 2);
        uint8_t* it = mem.get<uint8_t>((size_t)n * 2);
        uint8_t* ac = mem.get<uint8_t>((size_t)n);
        float* se = mem.get<float>((size_t)n * 2);
        double* pd = mem.get<double>((size_t)n);
        const MlpDev M{ s->mlp_frame, s->mlp_h1pre, static_cast<const u32x2_t*>(s->mlp_wfrag), s->mlp_params };
        hipLaunchKernelGGL(k_mlp_ab, dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, s->dev, M, s->n_patches, seed,
                           first, n, st, it, ac, se, pd);
        check(hipGetLastError(), "k_mlp_ab");
        check(hipMemcpyAsync(status, st, n * 2, hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(iters, it, n * 2, hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(accepted, ac, n, hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(seed_err, se, n * 2 * sizeof(float), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipMemcpyAsync(point_diff, pd, n * sizeof(double), hipMemcpyDeviceToHost, s->stream), "D2H");
        check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

}  // extern "C"
