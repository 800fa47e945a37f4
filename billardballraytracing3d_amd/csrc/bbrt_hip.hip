// bbrt_hip.hip -- gfx950 kernels and C ABI of the billiard bounce loop.
//
// One lane per live ray, one kernel launch per bounce; live rays are
// compacted between bounces with a wave ballot + prefix popcount and one
// atomic per wave.  Within a bounce a lane runs, in the reference's exact
// order and FP64 arithmetic:
//   BVH::intersect            include/BVH.hpp:75-128 (ray_box :14-37)
//   SurfacePatch::intersect   include/ScatObjects/SurfacePatch.hpp:253-303
//     stage 1: Moller-Trumbore  include/Utils/MollerTrumbore.hpp:32-64
//     triangle_uv_centroid      SurfacePatch.hpp:499-510
//     newton_refine             SurfacePatch.hpp:593-660 (+ optional polish)
//   back-face check, reflect, nudge   include/RayTracingDriver.hpp:206-237
//
// Stage 1 visits only the coarse cells whose FP32 box (all three offset
// copies of both triangles, widened by a margin that exceeds every rounding
// error of the box test and of Moller-Trumbore's acceptance) the ray enters,
// through a 3x3 -> 4x4 -> 3x3 -> 3x3 cell pyramid.  A cell outside its box
// cannot produce an accepted triangle, so the winner -- strictly smallest t,
// ties to the first in the reference's (i, j, sign, triangle) loop order --
// is the brute-force winner.  Ties are resolved by comparing that loop key.
//
// All FP64 decision arithmetic is compiled with -ffp-contract=off: no FMA
// contraction, IEEE division and square root, so results are bit-identical to
// the reference expressions evaluated in the same order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "scene_host.h"

using bbrt::BvhNode;
using bbrt::PatchMeta;

namespace {

thread_local std::string g_last_error;

constexpr int kStackCap = 48;
constexpr int kDefaultWavesPerEu = 3;         // persistent-kernel occupancy target (best of 2/3/4 on MI355X)
constexpr double kMtEps = 1e-9;               // MollerTrumbore.hpp:33
constexpr double kOnNodeTol = 1.0e-12;        // PatchInterpolation.hpp:21
constexpr int kNewtonMaxIters = 20;           // SurfacePatch.hpp:121
constexpr double kNewtonTol = 1e-3;           // SurfacePatch.hpp:122
constexpr double kDomainTol = 1e-6;           // SurfacePatch.hpp:639
constexpr double kDetTol = 1e-14;             // SurfacePatch.hpp:567
constexpr double kSelfEpsFactor = 1e-9;       // RayTracingDriver.hpp:328
constexpr double kSagittaClearance = 5.0;     // RayTracingDriver.hpp:337
constexpr double kBackfaceTol = 1e-9;         // RayTracingDriver.hpp:344
constexpr int kPolishMax = 8;                 // polish of the converged mode
constexpr double kPolishRelTol = 1e-13;

struct DevScene {
    const PatchMeta* meta;
    const double* nodes;
    const double* fields;
    const double* cxyz;
    const double* cell;
    const float* box_cell;
    const float* box_tile;
    const float* box_block;
    const float* box_super;
    const float* box_patch;
    const BvhNode* bvh;
    const int32_t* leaf;
    double diag;
    double cull_limit;   // culling is used while |origin|_inf <= cull_limit
};

struct RayState {
    double* o;   // [n][3]
    double* d;   // [n][3]
};

struct Outs {
    int8_t* status;
    int8_t* err_kind;
    int32_t* bounces;
    int32_t* err_patch;
    int32_t* hit_patch;
    double* hit_xyz;
    uint8_t* newton_iters;
    double* final_od;
    unsigned long long* hit_count;
    uint32_t* ray_work;
};

struct V3 { double x, y, z; };

__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
__device__ __forceinline__ double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 unit3(V3 v) {
    const double len = sqrt(dot3(v, v));
    return { v.x / len, v.y / len, v.z / len };
}
// std::max(a, b) / std::min(a, b) of the reference: a NaN second argument is ignored.
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

// Per-lane work counts (32-bit: one lane's share of one trace fits easily);
// reduced over the wave and added to the 64-bit totals once per kernel.
struct Counters {
    uint32_t bvh_tests, candidates, cull_tests, mt_tests, newton_calls, newton_iters, newton_flops, polish_fail,
        bounces;
};

// MollerTrumbore.hpp:32-64 (t only; the point and normal are unused by SurfacePatch).
__device__ __forceinline__ bool mt_t(const V3& o, const V3& d, const V3& v0, const V3& v1, const V3& v2, double& t_out) {
    const V3 e1 = sub3(v1, v0);
    const V3 e2 = sub3(v2, v0);
    const V3 pv = cross3(d, e2);
    const double det = dot3(e1, pv);
    if (fabs(det) < kMtEps) return false;
    const double inv = 1.0 / det;
    const V3 tv = sub3(o, v0);
    const double u = dot3(tv, pv) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const V3 qv = cross3(tv, e1);
    const double v = dot3(d, qv) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    const double t = dot3(e2, qv) * inv;
    if (t < kMtEps) return false;
    t_out = t;
    return true;
}

// BVH.hpp:14-37
__device__ __forceinline__ bool ray_box64(const V3& o, const double inv[3], const BvhNode& nd, double best_t) {
    double tmin = -INFINITY, tmax = INFINITY;
    const double oo[3] = { o.x, o.y, o.z };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t1 = (nd.lo[a] - oo[a]) * inv[a];
        double t2 = (nd.hi[a] - oo[a]) * inv[a];
        if (inv[a] < 0.0) { const double s = t1; t1 = t2; t2 = s; }
        tmin = smax(tmin, t1);
        tmax = smin(tmax, t2);
        if (tmax < tmin) return false;
    }
    if (tmin > best_t) return false;
    return tmax >= smax(tmin, 0.0);
}

// Conservative FP32 slab test against a widened box (no best-t bound).
struct Ray32 { float ox, oy, oz, ix, iy, iz; };
__device__ __forceinline__ bool box32(const float* __restrict__ b, const Ray32& r) {
    float t1 = (b[0] - r.ox) * r.ix, t2 = (b[3] - r.ox) * r.ix;
    if (r.ix < 0.0f) { const float s = t1; t1 = t2; t2 = s; }
    float tmin = -INFINITY, tmax = INFINITY;
    tmin = (tmin < t1) ? t1 : tmin; tmax = (t2 < tmax) ? t2 : tmax;
    t1 = (b[1] - r.oy) * r.iy; t2 = (b[4] - r.oy) * r.iy;
    if (r.iy < 0.0f) { const float s = t1; t1 = t2; t2 = s; }
    tmin = (tmin < t1) ? t1 : tmin; tmax = (t2 < tmax) ? t2 : tmax;
    t1 = (b[2] - r.oz) * r.iz; t2 = (b[5] - r.oz) * r.iz;
    if (r.iz < 0.0f) { const float s = t1; t1 = t2; t2 = s; }
    tmin = (tmin < t1) ? t1 : tmin; tmax = (t2 < tmax) ? t2 : tmax;
    return tmax >= tmin && tmax >= 0.0f;
}

struct Stage1 {
    double t;
    int key;      // ((i*99 + j)*3 + sign index)*2 + triangle: the reference's loop position
    bool have;
};

// The six Moller-Trumbore tests of coarse cell (i, j) (SurfacePatch.hpp:274-294).
__device__ __forceinline__ void cell_tests(const double* __restrict__ G, const double* __restrict__ C, int i, int j,
                                           const V3& o, const V3& d, Stage1& best, Counters& c) {
    const double* a = G + ((size_t)i * bbrt::kCoarseN + j) * 3;
    const double* b = G + ((size_t)(i + 1) * bbrt::kCoarseN + j) * 3;
    const V3 p00 = { a[0], a[1], a[2] }, p01 = { a[3], a[4], a[5] };
    const V3 p10 = { b[0], b[1], b[2] }, p11 = { b[3], b[4], b[5] };
    const double* cc = C + ((size_t)i * bbrt::kCells + j) * 4;
    const double s = cc[0];
    const V3 cn = { cc[1], cc[2], cc[3] };
    const int base = (i * bbrt::kCells + j) * 6;
    const double signs[3] = { 0.0, 1.0, -1.0 };
#pragma unroll
    for (int si = 0; si < 3; ++si) {
        const double dd = signs[si] * s;
        const V3 q00 = { p00.x + cn.x * dd, p00.y + cn.y * dd, p00.z + cn.z * dd };
        const V3 q10 = { p10.x + cn.x * dd, p10.y + cn.y * dd, p10.z + cn.z * dd };
        const V3 q01 = { p01.x + cn.x * dd, p01.y + cn.y * dd, p01.z + cn.z * dd };
        const V3 q11 = { p11.x + cn.x * dd, p11.y + cn.y * dd, p11.z + cn.z * dd };
        double t;
        int key = base + 2 * si;
        if (mt_t(o, d, q00, q10, q11, t) && (!best.have || t < best.t || (t == best.t && key < best.key))) {
            best.have = true; best.t = t; best.key = key;
        }
        key += 1;
        if (mt_t(o, d, q00, q11, q01, t) && (!best.have || t < best.t || (t == best.t && key < best.key))) {
            best.have = true; best.t = t; best.key = key;
        }
    }
    c.mt_tests += 6;
}

// Stage 1 over the whole coarse grid in the reference's order (brute force).
__device__ void stage1_brute(const double* G, const double* C, const V3& o, const V3& d, Stage1& best, Counters& c) {
    for (int i = 0; i < bbrt::kCells; ++i)
        for (int j = 0; j < bbrt::kCells; ++j) cell_tests(G, C, i, j, o, d, best, c);
}

// Loads one aligned group of NB boxes (NF floats) and returns the mask of the
// boxes the ray enters.
template <int NB, int NF>
__device__ __forceinline__ uint32_t group_mask(const float* __restrict__ g, const Ray32& r) {
    float v[NF];
    const float4* g4 = reinterpret_cast<const float4*>(g);
#pragma unroll
    for (int q = 0; q < NF / 4; ++q) {
        const float4 x = g4[q];
        v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
    }
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < NB; ++q)
        if (box32(v + 6 * q, r)) m |= 1u << q;
    return m;
}

// Stage 1 through the culling pyramid (grouped layout: one load per level).
__device__ void stage1_culled(const DevScene& S, int64_t k, const double* G, const double* C, const V3& o,
                              const V3& d, const Ray32& r, Stage1& best, Counters& c) {
    const float* gs = S.box_super + (size_t)k * bbrt::kSuperFloats;
    const float* gb = S.box_block + (size_t)k * bbrt::kBlockFloats;
    const float* gt = S.box_tile + (size_t)k * bbrt::kTileFloats;
    const float* gc = S.box_cell + (size_t)k * bbrt::kCellFloats;
    uint32_t ms = group_mask<9, bbrt::kG9>(gs, r);
    c.cull_tests += 9;
    while (ms) {
        const int sq = __ffs(ms) - 1;
        ms &= ms - 1;
        const int si = sq / 3, sj = sq % 3;
        uint32_t mb = group_mask<16, bbrt::kG16>(gb + sq * bbrt::kG16, r);
        c.cull_tests += 16;
        while (mb) {
            const int bq = __ffs(mb) - 1;
            mb &= mb - 1;
            const int bi = bbrt::kBlocksPerSuper * si + (bq >> 2), bj = bbrt::kBlocksPerSuper * sj + (bq & 3);
            uint32_t mt = group_mask<9, bbrt::kG9>(gt + (bi * bbrt::kBlockN + bj) * bbrt::kG9, r);
            c.cull_tests += 9;
            while (mt) {
                const int tq = __ffs(mt) - 1;
                mt &= mt - 1;
                const int ti = 3 * bi + tq / 3, tj = 3 * bj + tq % 3;
                uint32_t mc = group_mask<9, bbrt::kG9>(gc + (ti * bbrt::kTileN + tj) * bbrt::kG9, r);
                c.cull_tests += 9;
                while (mc) {
                    const int cq = __ffs(mc) - 1;
                    mc &= mc - 1;
                    cell_tests(G, C, 3 * ti + cq / 3, 3 * tj + cq % 3, o, d, best, c);
                }
            }
        }
    }
}

struct PHit {
    int status;   // 1 hit, 0 miss, -kind error
    double t;
    V3 p, n;
    int iters;
};

// SurfacePatch.hpp:518-538
template <int N>
__device__ __forceinline__ void basis(const double* __restrict__ nodes, const double* __restrict__ w, int n, double x,
                                      double* b) {
    constexpr int M = N > 0 ? N : bbrt::kMaxNodes1D;
    const int cnt = N > 0 ? N : n;
    int hot = -1;
#pragma unroll
    for (int i = 0; i < M; ++i)
        if (i < cnt && hot < 0 && fabs(x - nodes[i]) < kOnNodeTol) hot = i;
    if (hot >= 0) {
#pragma unroll
        for (int i = 0; i < M; ++i) b[i] = (i == hot) ? 1.0 : 0.0;
        return;
    }
    double denom = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        if (i < cnt) {
            const double li = w[i] / (x - nodes[i]);
            b[i] = li;
            denom += li;
        } else {
            b[i] = 0.0;
        }
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
        if (i < cnt) b[i] /= denom;
}

// SurfacePatch.hpp:542-555.  For a fixed node count all values are loaded
// first (one round trip), then combined in the reference's order, skipping
// rows whose u-basis value is zero exactly as the reference does.
template <int N>
__device__ __forceinline__ double field(const double* lu, const double* lv, const double* __restrict__ f, int im,
                                        int jm) {
    if constexpr (N > 0) {
        // keep the scheduler from hoisting the next field's loads above this one
        __builtin_amdgcn_sched_barrier(0);
        double v[N * N];
#pragma unroll
        for (int k = 0; k < N * N; ++k) v[k] = f[k];
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double li = lu[i];
            if (li != 0.0) {
                double row = 0.0;
#pragma unroll
                for (int j = 0; j < N; ++j) row += lv[j] * v[i * N + j];
                sum += li * row;
            }
        }
        return sum;
    }
    constexpr int M = N > 0 ? N : bbrt::kMaxNodes1D;
    const int ci = N > 0 ? N : im, cj = N > 0 ? N : jm;
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        if (i < ci) {
            const double li = lu[i];
            if (li != 0.0) {
                double row = 0.0;
#pragma unroll
                for (int j = 0; j < M; ++j)
                    if (j < cj) row += lv[j] * f[i * cj + j];
                sum += li * row;
            }
        }
    }
    return sum;
}

// SurfacePatch.hpp:559-579
__device__ __forceinline__ bool solve3(const double J[3][3], const V3& b, double x[3]) {
    const double a11 = J[0][0], a12 = J[0][1], a13 = J[0][2];
    const double a21 = J[1][0], a22 = J[1][1], a23 = J[1][2];
    const double a31 = J[2][0], a32 = J[2][1], a33 = J[2][2];
    const double det = a11 * (a22 * a33 - a23 * a32) - a12 * (a21 * a33 - a23 * a31) + a13 * (a21 * a32 - a22 * a31);
    if (fabs(det) < kDetTol) return false;
    const double bx = b.x, by = b.y, bz = b.z;
    const double det0 = bx * (a22 * a33 - a23 * a32) - a12 * (by * a33 - a23 * bz) + a13 * (by * a32 - a22 * bz);
    const double det1 = a11 * (by * a33 - a23 * bz) - bx * (a21 * a33 - a23 * a31) + a13 * (a21 * bz - by * a31);
    const double det2 = a11 * (a22 * bz - by * a32) - a12 * (a21 * bz - by * a31) + bx * (a21 * a32 - a22 * a31);
    x[0] = det0 / det; x[1] = det1 / det; x[2] = det2 / det;
    return true;
}

// Newton on (u, v, t), SurfacePatch.hpp:593-660; mode 1 adds the polish.
template <int N>
__device__ PHit newton(const DevScene& S, const PatchMeta& M, const V3& o, const V3& d, double u, double v, double t,
                       int mode, Counters& c) {
    constexpr int L = N > 0 ? N : bbrt::kMaxNodes1D;
    const int im = N > 0 ? N : M.imax, jm = N > 0 ? N : M.jmax;
    const double* un = S.nodes + M.node_off;
    const double* uw = un + im;
    const double* vn = uw + im;
    const double* vw = vn + jm;
    const double* F = S.fields + M.field_off;
    const int nn = im * jm;
    double lu[L], lv[L];
    PHit r{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    c.newton_calls++;
    auto step = [&](double delta[3]) -> bool {
        basis<N>(un, uw, im, u, lu);
        basis<N>(vn, vw, jm, v, lv);
        const V3 X = { field<N>(lu, lv, F, im, jm), field<N>(lu, lv, F + nn, im, jm),
                       field<N>(lu, lv, F + 2 * nn, im, jm) };
        const V3 Xu = { field<N>(lu, lv, F + 3 * nn, im, jm), field<N>(lu, lv, F + 4 * nn, im, jm),
                        field<N>(lu, lv, F + 5 * nn, im, jm) };
        const V3 Xv = { field<N>(lu, lv, F + 6 * nn, im, jm), field<N>(lu, lv, F + 7 * nn, im, jm),
                        field<N>(lu, lv, F + 8 * nn, im, jm) };
        const V3 R = { o.x + t * d.x, o.y + t * d.y, o.z + t * d.z };
        const V3 Fv = sub3(X, R);
        const double J[3][3] = { { Xu.x, Xv.x, -d.x }, { Xu.y, Xv.y, -d.y }, { Xu.z, Xv.z, -d.z } };
        return solve3(J, Fv, delta);
    };
    // flops per iteration and per final evaluation (N = nodes per side)
    const uint32_t Ni = (uint32_t)im;
    const uint32_t I_it = 8 * Ni + 18 * Ni * (Ni + 1) + 77;
    const uint32_t F_fin = 8 * Ni + 12 * Ni * (Ni + 1) + 9;
    int iters = 0;
    double step_norm = INFINITY;
    for (int it = 0; it < kNewtonMaxIters; ++it) {
        double delta[3];
        if (!step(delta)) {
            c.newton_iters += iters; c.newton_flops += iters * I_it;
            r.status = -bbrt::kErrSingular; r.iters = iters; return r;
        }
        u -= delta[0]; v -= delta[1]; t -= delta[2];
        ++iters;
        step_norm = sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
        if (step_norm < kNewtonTol) break;
    }
    if (step_norm >= kNewtonTol) {
        c.newton_iters += iters; c.newton_flops += iters * I_it;
        r.status = -bbrt::kErrNoConv; r.iters = iters; return r;
    }
    if (u < -1.0 - kDomainTol || u > 1.0 + kDomainTol || v < -1.0 - kDomainTol || v > 1.0 + kDomainTol) {
        c.newton_iters += iters; c.newton_flops += iters * I_it;
        r.status = -bbrt::kErrDomain; r.iters = iters; return r;
    }
    if (mode == 1) {
        bool ok = false;
        for (int k = 0; k < kPolishMax; ++k) {
            double delta[3];
            if (!step(delta)) break;
            u -= delta[0]; v -= delta[1]; t -= delta[2];
            ++iters;
            const double sn = sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
            if (sn < kPolishRelTol * smax(1.0, fabs(t))) { ok = true; break; }
        }
        if (!ok) c.polish_fail++;
    }
    c.newton_iters += iters; c.newton_flops += iters * I_it + F_fin;
    basis<N>(un, uw, im, u, lu);
    basis<N>(vn, vw, jm, v, lv);
    r.status = 1; r.t = t; r.iters = iters;
    r.p = { field<N>(lu, lv, F, im, jm), field<N>(lu, lv, F + nn, im, jm), field<N>(lu, lv, F + 2 * nn, im, jm) };
    r.n = unit3({ field<N>(lu, lv, F + 9 * nn, im, jm), field<N>(lu, lv, F + 10 * nn, im, jm),
                  field<N>(lu, lv, F + 11 * nn, im, jm) });
    return r;
}

// SurfacePatch::intersect (SurfacePatch.hpp:253-303)
__device__ PHit patch_intersect(const DevScene& S, int64_t k, const V3& o, const V3& d, const Ray32& r32, bool cull,
                                int mode, Counters& c) {
    const PatchMeta M = S.meta[k];
    const double* G = S.cxyz + (size_t)k * bbrt::kCoarseN * bbrt::kCoarseN * 3;
    const double* C = S.cell + (size_t)k * bbrt::kCells * bbrt::kCells * 4;
    Stage1 best{ 0.0, 0, false };
    c.candidates++;
    if (cull && M.cull) stage1_culled(S, k, G, C, o, d, r32, best, c);
    else stage1_brute(G, C, o, d, best, c);
    if (!best.have) return PHit{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    const int cellk = best.key / 6;
    const int i = cellk / bbrt::kCells, j = cellk % bbrt::kCells;
    const bool second = (best.key & 1) != 0;
    // coarse_.u[i] = -1 + 2 i / 99 (SurfacePatch.hpp:421), then triangle_uv_centroid (:499-510)
    const double ui = -1.0 + 2.0 * static_cast<double>(i) / (bbrt::kCoarseN - 1);
    const double uip1 = -1.0 + 2.0 * static_cast<double>(i + 1) / (bbrt::kCoarseN - 1);
    const double vj = -1.0 + 2.0 * static_cast<double>(j) / (bbrt::kCoarseN - 1);
    const double vjp1 = -1.0 + 2.0 * static_cast<double>(j + 1) / (bbrt::kCoarseN - 1);
    double u0, v0;
    if (!second) { u0 = (ui + uip1 + uip1) / 3.0; v0 = (vj + vj + vjp1) / 3.0; }
    else { u0 = (ui + ui + uip1) / 3.0; v0 = (vj + vjp1 + vjp1) / 3.0; }
    if (M.imax == 3 && M.jmax == 3) return newton<3>(S, M, o, d, u0, v0, best.t, mode, c);
    if (M.imax == 5 && M.jmax == 5) return newton<5>(S, M, o, d, u0, v0, best.t, mode, c);
    return newton<0>(S, M, o, d, u0, v0, best.t, mode, c);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void flush_counters(const Counters& c, unsigned long long* ctr) {
    const unsigned long long v[9] = { c.bvh_tests, c.candidates, c.cull_tests, c.mt_tests, c.newton_calls,
                                      c.newton_iters, c.newton_flops, c.polish_fail, c.bounces };  // widened
    const int slot[9] = { bbrt::kCtrBvh, bbrt::kCtrCand, bbrt::kCtrCull, bbrt::kCtrMt, bbrt::kCtrNewtonCalls,
                          bbrt::kCtrNewtonIters, bbrt::kCtrNewtonFlops, bbrt::kCtrPolishFail, bbrt::kCtrBounces };
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const unsigned long long s = wave_sum(v[q]);
        if (__lane_id() == 0 && s) atomicAdd(ctr + slot[q], s);
    }
}

__device__ __forceinline__ void append_live(bool alive, int32_t r, int32_t* live_out, int32_t* n_out) {
    const unsigned long long mask = __ballot(alive);
    const int lane = __lane_id();
    int base = 0;
    if (lane == 0 && mask) base = atomicAdd(n_out, (int)__popcll(mask));
    base = __shfl(base, 0);
    if (alive) {
        const unsigned long long below = mask & ((1ull << lane) - 1ull);
        live_out[base + __popcll(below)] = r;
    }
}

__device__ __forceinline__ void write_final(const Outs& O, int32_t r, const V3& o, const V3& d) {
    if (O.final_od) {
        double* f = O.final_od + (size_t)r * 6;
        f[0] = o.x; f[1] = o.y; f[2] = o.z; f[3] = d.x; f[4] = d.y; f[5] = d.z;
    }
}

// The part of ray_box (BVH.hpp:14-37) that does not depend on best_t: false
// when the box is rejected whatever best_t is, else the entry distance.
__device__ __forceinline__ bool box_static(const V3& o, const double inv[3], const BvhNode& nd, double& tmin_out) {
    double tmin = -INFINITY, tmax = INFINITY;
    const double oo[3] = { o.x, o.y, o.z };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t1 = (nd.lo[a] - oo[a]) * inv[a];
        double t2 = (nd.hi[a] - oo[a]) * inv[a];
        if (inv[a] < 0.0) { const double s = t1; t1 = t2; t2 = s; }
        tmin = smax(tmin, t1);
        tmax = smin(tmax, t2);
        if (tmax < tmin) return false;
    }
    tmin_out = tmin;
    return tmax >= smax(tmin, 0.0);
}

// Traversal stack entry: idx < 0 marks a node whose box test already failed;
// meta >= 0 is an internal node (right child << 2 | split axis), meta < 0 a
// leaf (-1 - (start << 3 | count)); tmin is the node's entry distance.
struct Ent {
    int32_t idx;
    int32_t meta;
    double tmin;
};

__device__ __forceinline__ Ent make_ent(int32_t idx, const BvhNode& nd, bool ok, double tmin) {
    Ent e;
    e.idx = ok ? idx : -1;
    e.meta = (nd.right >= 0) ? ((nd.right << 2) | nd.axis) : (-1 - ((nd.start << 3) | nd.count));
    e.tmin = tmin;
    return e;
}

// BVH::intersect (BVH.hpp:75-128).  A node's box is tested when its parent is
// expanded: the best_t-independent part of ray_box then, the `tmin > best_t`
// prune when the node is popped -- the same outcome as testing it whole at the
// pop, since best_t only decreases.  A child that fails is still pushed (as a
// dead entry) so the visit order, the stack depth and the node-test count are
// the reference's.  The near child stays in registers instead of being pushed
// and popped at once.  Returns 0 or an error kind.
__device__ int traverse(const DevScene& S, const V3& o, const V3& d, const Ray32& r32, bool cull, int mode,
                        bool root_live, PHit& best, int32_t& best_obj, int32_t& err_obj, Counters& c) {
    const double inv[3] = { 1.0 / d.x, 1.0 / d.y, 1.0 / d.z };
    const bool neg[3] = { inv[0] < 0.0, inv[1] < 0.0, inv[2] < 0.0 };
    double best_t = INFINITY;
    best_obj = -1;
    err_obj = -1;
    Ent stack[kStackCap];
    int sp = 0;
    Ent cur;
    {
        const BvhNode root = S.bvh[0];
        double tm = -INFINITY;
        const bool ok = root_live ? true : box_static(o, inv, root, tm);
        cur = make_ent(0, root, ok, tm);
    }
    // While-while form: the inner loop walks internal nodes (cheap) until the
    // lane stands on a leaf to intersect or runs out of nodes; leaves are then
    // intersected with the whole wave in step, instead of every lane paying
    // for a patch intersection whenever any lane reaches a leaf.
    for (;;) {
        bool at_leaf = false;
        for (;;) {
            c.bvh_tests++;
            const bool live = cur.idx >= 0 && !(cur.tmin > best_t);
            if (live && cur.meta < 0) { at_leaf = true; break; }
            if (!live) {
                if (sp == 0) break;
                cur = stack[--sp];
                continue;
            }
            const int32_t left = cur.idx + 1, right = cur.meta >> 2;
            const BvhNode nl = S.bvh[left];
            const BvhNode nr = S.bvh[right];
            double tl = 0.0, tr = 0.0;
            const bool okl = box_static(o, inv, nl, tl);
            const bool okr = box_static(o, inv, nr, tr);
            if (sp + 1 > kStackCap) return bbrt::kErrStack;
            const Ent el = make_ent(left, nl, okl, tl), er = make_ent(right, nr, okr, tr);
            if (neg[cur.meta & 3]) { stack[sp++] = el; cur = er; }
            else { stack[sp++] = er; cur = el; }
        }
        if (!at_leaf) break;
        const int v = -1 - cur.meta;
        const int start = v >> 3, count = v & 7;
        for (int q0 = 0; q0 < count; q0 += 4) {
            // Batch prefilter: the conservative box of each object's whole
            // culling pyramid, loaded together.  An object whose box the ray
            // does not enter has no stage-1 hit: a miss, with no further loads.
            int32_t ids[4];
            uint32_t enter = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) ids[q] = (q0 + q < count) ? S.leaf[start + q0 + q] : -1;
            if (cull) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (ids[q] >= 0) {
                        const float4* b4 = reinterpret_cast<const float4*>(S.box_patch + (size_t)ids[q] * 8);
                        const float4 a = b4[0], b = b4[1];
                        const float bx[6] = { a.x, a.y, a.z, a.w, b.x, b.y };
                        if (box32(bx, r32)) enter |= 1u << q;
                    }
                c.cull_tests += min(4, count - q0);
            } else {
                enter = 0xFu;
            }
            for (int q = 0; q < 4 && q0 + q < count; ++q) {
                const int32_t obj = ids[q];
                if (!(enter & (1u << q))) { c.candidates++; continue; }
                const PHit h = patch_intersect(S, obj, o, d, r32, cull, mode, c);
                if (h.status < 0) { err_obj = obj; return -h.status; }
                if (h.status == 1 && h.t < best_t) { best_t = h.t; best = h; best_obj = obj; }
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return 0;
}

// One bounce of ray r: RayTracingDriver.hpp:186-237.  Returns true while the
// ray continues; o and d are advanced in place.
// The end of one bounce, RayTracingDriver.hpp:190-237: error, escape, back
// face, or record + reflect + nudge.  Returns true while the ray continues;
// o and d are advanced in place.
__device__ bool finish_bounce(const DevScene& S, const Outs& O, int32_t r, V3& o, V3& d, int bounce, int max_bounces,
                              int record_k, uint32_t flags, int err, int32_t err_obj, const PHit& best,
                              int32_t best_obj, Counters& c) {
    if (err) {
        if (O.status) O.status[r] = bbrt::kError;
        if (O.err_kind) O.err_kind[r] = (int8_t)err;
        if (O.err_patch) O.err_patch[r] = err_obj;
        if (O.bounces) O.bounces[r] = bounce;
        write_final(O, r, o, d);
        return false;
    }
    if (best_obj < 0) {
        if (O.status) O.status[r] = bbrt::kEscaped;
        if (O.bounces) O.bounces[r] = bounce;
        write_final(O, r, o, d);
        return false;
    }
    if (dot3(d, best.n) > kBackfaceTol) {
        if (O.status) O.status[r] = bbrt::kError;
        if (O.err_kind) O.err_kind[r] = bbrt::kErrBackface;
        if (O.err_patch) O.err_patch[r] = best_obj;
        if (O.bounces) O.bounces[r] = bounce;
        write_final(O, r, o, d);
        return false;
    }
    if (bounce < record_k) {
        const size_t at = (size_t)r * record_k + bounce;
        if (O.hit_patch) O.hit_patch[at] = best_obj;
        if (O.newton_iters) O.newton_iters[at] = (uint8_t)best.iters;
        if (O.hit_xyz) { O.hit_xyz[3 * at] = best.p.x; O.hit_xyz[3 * at + 1] = best.p.y; O.hit_xyz[3 * at + 2] = best.p.z; }
    }
    if (O.hit_count && (bounce >= 1 || (flags & BBRT_FLAG_COUNT_DIRECT_HIT)))
        atomicAdd(O.hit_count + best_obj, 1ull);
    c.bounces++;
    // reflect (RayTracingDriver.hpp:405-408) and nudge (:234-237, :260-265)
    const V3& nn = best.n;
    const double dn = d.x * nn.x + d.y * nn.y + d.z * nn.z;
    d = { d.x - 2.0 * dn * nn.x, d.y - 2.0 * dn * nn.y, d.z - 2.0 * dn * nn.z };
    const double eps = smax(kSelfEpsFactor * S.diag, kSagittaClearance * S.meta[best_obj].max_sag);
    o = { best.p.x + d.x * eps, best.p.y + d.y * eps, best.p.z + d.z * eps };
    if (bounce + 1 >= max_bounces) {
        if (O.status) O.status[r] = bbrt::kTrapped;
        if (O.bounces) O.bounces[r] = bounce + 1;
        write_final(O, r, o, d);
        return false;
    }
    return true;
}

__device__ __forceinline__ bool cull_ok(const DevScene& S, const V3& o, const V3& d, uint32_t flags) {
    const bool finite = isfinite(o.x) && isfinite(o.y) && isfinite(o.z) && isfinite(d.x) && isfinite(d.y) &&
                        isfinite(d.z);
    return !(flags & BBRT_FLAG_NO_CULL) && finite && fabs(o.x) <= S.cull_limit && fabs(o.y) <= S.cull_limit &&
           fabs(o.z) <= S.cull_limit;
}

__device__ __forceinline__ void add_ray_work(const Outs& O, int32_t r, const Counters& c, const Counters& c0) {
    uint32_t* w = O.ray_work + 4 * (size_t)r;
    w[0] += (uint32_t)(c.bvh_tests - c0.bvh_tests);
    w[1] += (uint32_t)(c.cull_tests - c0.cull_tests);
    w[2] += (uint32_t)(c.mt_tests - c0.mt_tests);
    w[3] += (uint32_t)(c.newton_iters - c0.newton_iters);
}

// One whole bounce of ray r (the wavefront schedule).
template <bool WORK>
__device__ bool bounce_once(const DevScene& S, const Outs& O, int32_t r, V3& o, V3& d, int bounce, int max_bounces,
                            int record_k, int mode, uint32_t flags, bool root_live, Counters& c) {
    Counters c0;
    if constexpr (WORK) c0 = c;
    const bool cull = cull_ok(S, o, d, flags);
    const Ray32 r32 = { (float)o.x, (float)o.y, (float)o.z, 1.0f / (float)d.x, 1.0f / (float)d.y, 1.0f / (float)d.z };
    PHit best{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    int32_t best_obj, err_obj;
    const int err = traverse(S, o, d, r32, cull, mode, root_live, best, best_obj, err_obj, c);
    const bool alive = finish_bounce(S, O, r, o, d, bounce, max_bounces, record_k, flags, err, err_obj, best, best_obj, c);
    if constexpr (WORK) add_ray_work(O, r, c, c0);
    return alive;
}

// RayTracingDriver::trace_ray, lines 172-184: zero-direction check and
// normalize_dir (:398-401); then the root box test of the first BVH query
// (BVH.hpp:91-97 with best_t = inf), so rays that miss the whole scene leave
// here as Escaped after one node test.  The rest enter the live list.
__global__ void k_init(const double* __restrict__ origins, const double* __restrict__ dirs, int64_t n,
                       const BvhNode* __restrict__ bvh, RayState R, Outs O, int max_bounces, int32_t* live_out,
                       int32_t* n_out, unsigned long long* ctr) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    int32_t r = (int32_t)g;
    unsigned long long root_tests = 0;
    if (g < n) {
        const V3 o = { origins[3 * g], origins[3 * g + 1], origins[3 * g + 2] };
        const V3 dr = { dirs[3 * g], dirs[3 * g + 1], dirs[3 * g + 2] };
        if (O.ray_work) { O.ray_work[4 * g] = 0; O.ray_work[4 * g + 1] = 0; O.ray_work[4 * g + 2] = 0; O.ray_work[4 * g + 3] = 0; }
        if (O.err_kind) O.err_kind[g] = bbrt::kErrNone;
        if (O.err_patch) O.err_patch[g] = -1;
        const double len2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
        if (len2 < 1e-300) {
            if (O.status) O.status[g] = bbrt::kError;
            if (O.err_kind) O.err_kind[g] = bbrt::kErrZeroDir;
            if (O.bounces) O.bounces[g] = 0;
            write_final(O, r, o, dr);
        } else {
            const double len = sqrt(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
            const V3 d = { dr.x / len, dr.y / len, dr.z / len };
            const double inv[3] = { 1.0 / d.x, 1.0 / d.y, 1.0 / d.z };
            double tm;
            if (max_bounces <= 0) {
                if (O.status) O.status[g] = bbrt::kTrapped;
                if (O.bounces) O.bounces[g] = 0;
                write_final(O, r, o, d);
            } else if (!box_static(o, inv, bvh[0], tm)) {
                root_tests = 1;
                if (O.status) O.status[g] = bbrt::kEscaped;
                if (O.bounces) O.bounces[g] = 0;
                if (O.ray_work) O.ray_work[4 * g] = 1;
                write_final(O, r, o, d);
            } else {
                R.o[3 * g] = o.x; R.o[3 * g + 1] = o.y; R.o[3 * g + 2] = o.z;
                R.d[3 * g] = d.x; R.d[3 * g + 1] = d.y; R.d[3 * g + 2] = d.z;
                alive = true;
            }
        }
    }
    append_live(alive, r, live_out, n_out);
    const unsigned long long in_range = __ballot(g < n);
    const unsigned long long rt = wave_sum(root_tests);
    if (__lane_id() == 0) {
        if (in_range) atomicAdd(ctr + bbrt::kCtrRays, (unsigned long long)__popcll(in_range));
        if (rt) atomicAdd(ctr + bbrt::kCtrBvh, rt);
    }
}

// Persistent bounce loop as a per-lane state machine under a wave-uniform
// outer loop.  Each iteration: lanes without a ray take one from the live list
// (one wave-aggregated atomic); traversing lanes walk internal nodes until they
// stand on a leaf (while-while); lanes on a leaf intersect one of its objects;
// lanes whose bounce is complete record it, reflect, and start the next
// bounce at once -- a lane never waits for another lane's whole bounce.  Every
// ray ends within max_bounces bounces and the list is finite, so every wave
// leaves the loop.
enum : int { kPhTrav = 1, kPhLeaf = 2, kPhDone = 3 };

template <int W, bool WORK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W)))
k_persist(DevScene S, RayState R, const int32_t* __restrict__ live, const int32_t* __restrict__ n_live_p,
          int32_t* head, Outs O, int max_bounces, int record_k, int mode, uint32_t flags, unsigned long long* ctr) {
    const int n_live = *n_live_p;
    const int lane = __lane_id();
    Counters c = {};
    Counters c0 = {};
    int32_t r = -1;   // -1: wants a ray, -2: list exhausted
    int b = 0, phase = 0;
    V3 o = { 0, 0, 0 }, d = { 0, 0, 0 };
    double inv[3] = { 0, 0, 0 };
    bool cull = false;
    Ray32 r32 = { 0, 0, 0, 0, 0, 0 };
    Ent stack[kStackCap];
    int sp = 0;
    Ent cur = { -1, 0, 0.0 };
    double best_t = INFINITY;
    PHit best{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    int32_t best_obj = -1, err_obj = -1;
    int err = 0;
    int leaf_start = 0, leaf_count = 0, q = 0;
    uint32_t enter = 0;

    // bounce set-up: RayTracingDriver.hpp:187 and the root of BVH.hpp:89
    auto start_bounce = [&](bool root_live) {
        cull = cull_ok(S, o, d, flags);
        r32 = { (float)o.x, (float)o.y, (float)o.z, 1.0f / (float)d.x, 1.0f / (float)d.y, 1.0f / (float)d.z };
        inv[0] = 1.0 / d.x; inv[1] = 1.0 / d.y; inv[2] = 1.0 / d.z;
        best_t = INFINITY;
        best_obj = -1; err_obj = -1; err = 0; sp = 0;
        const BvhNode root = S.bvh[0];
        double tm = -INFINITY;
        const bool ok = root_live ? true : box_static(o, inv, root, tm);
        cur = make_ent(0, root, ok, tm);
        phase = kPhTrav;
        if constexpr (WORK) c0 = c;
    };

    for (;;) {
        const unsigned long long want = __ballot(r == -1);
        if (want) {
            const int leader = __ffsll((long long)want) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(head, (int)__popcll(want));
            base = __shfl(base, leader);
            if (r == -1) {
                const int qq = base + (int)__popcll(want & ((1ull << lane) - 1ull));
                if (qq < n_live) {
                    r = live[qq];
                    o = { R.o[3 * r], R.o[3 * r + 1], R.o[3 * r + 2] };
                    d = { R.d[3 * r], R.d[3 * r + 1], R.d[3 * r + 2] };
                    b = 0;
                    start_bounce(true);   // k_init already tested the root box
                } else {
                    r = -2;
                }
            }
        }
        if (__ballot(r >= 0) == 0) break;

        // traversal: internal nodes until a leaf to intersect, or the end (BVH.hpp:91-126)
        if (r >= 0 && phase == kPhTrav) {
            for (;;) {
                c.bvh_tests++;
                const bool live_node = cur.idx >= 0 && !(cur.tmin > best_t);
                if (live_node && cur.meta < 0) {
                    const int v = -1 - cur.meta;
                    leaf_start = v >> 3; leaf_count = v & 7; q = 0;
                    // the conservative box of each object's culling pyramid, for the whole leaf at once
                    enter = 0;
                    if (cull) {
                        for (int k = 0; k < leaf_count; ++k) {
                            const int32_t obj = S.leaf[leaf_start + k];
                            const float4* b4 = reinterpret_cast<const float4*>(S.box_patch + (size_t)obj * 8);
                            const float4 a = b4[0], bb = b4[1];
                            const float bx[6] = { a.x, a.y, a.z, a.w, bb.x, bb.y };
                            if (box32(bx, r32)) enter |= 1u << k;
                        }
                        c.cull_tests += leaf_count;
                    } else {
                        enter = 0x7Fu;
                    }
                    phase = kPhLeaf;
                    break;
                }
                if (!live_node) {
                    if (sp == 0) { phase = kPhDone; break; }
                    cur = stack[--sp];
                    continue;
                }
                const int32_t left = cur.idx + 1, right = cur.meta >> 2;
                const BvhNode nl = S.bvh[left];
                const BvhNode nr = S.bvh[right];
                double tl = 0.0, tr = 0.0;
                const bool okl = box_static(o, inv, nl, tl);
                const bool okr = box_static(o, inv, nr, tr);
                if (sp + 1 > kStackCap) { err = bbrt::kErrStack; phase = kPhDone; break; }
                const Ent el = make_ent(left, nl, okl, tl), er = make_ent(right, nr, okr, tr);
                const bool negax = (cur.meta & 3) == 0 ? inv[0] < 0.0 : ((cur.meta & 3) == 1 ? inv[1] < 0.0 : inv[2] < 0.0);
                if (negax) { stack[sp++] = el; cur = er; }
                else { stack[sp++] = er; cur = el; }
            }
        }

        // one leaf object: SurfacePatch::intersect and the best-hit update (BVH.hpp:101-109)
        if (r >= 0 && phase == kPhLeaf) {
            const int32_t obj = S.leaf[leaf_start + q];
            if (enter & (1u << q)) {
                const PHit h = patch_intersect(S, obj, o, d, r32, cull, mode, c);
                if (h.status < 0) { err = -h.status; err_obj = obj; phase = kPhDone; }
                else if (h.status == 1 && h.t < best_t) { best_t = h.t; best = h; best_obj = obj; }
            } else {
                c.candidates++;
            }
            if (phase == kPhLeaf && ++q == leaf_count) {
                if (sp == 0) phase = kPhDone;
                else { cur = stack[--sp]; phase = kPhTrav; }
            }
        }

        if (r >= 0 && phase == kPhDone) {
            const bool alive = finish_bounce(S, O, r, o, d, b, max_bounces, record_k, flags, err, err_obj, best,
                                             best_obj, c);
            if constexpr (WORK) add_ray_work(O, r, c, c0);
            if (alive) { ++b; start_bounce(false); }
            else { r = -1; phase = 0; }
        }
    }
    flush_counters(c, ctr);
}

// One bounce of every live ray, compacting the survivors (the wavefront schedule).
__global__ void __launch_bounds__(256) k_bounce(DevScene S, RayState R, const int32_t* __restrict__ live_in,
                                                const int32_t* __restrict__ n_in, int32_t* live_out, int32_t* n_out,
                                                Outs O, int bounce, int max_bounces, int record_k, int mode,
                                                uint32_t flags, unsigned long long* ctr) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_live = *n_in;
    Counters c = {};
    bool alive = false;
    int32_t r = -1;
    if (g < n_live) {
        r = live_in[g];
        V3 o = { R.o[3 * r], R.o[3 * r + 1], R.o[3 * r + 2] };
        V3 d = { R.d[3 * r], R.d[3 * r + 1], R.d[3 * r + 2] };
        alive = O.ray_work ? bounce_once<true>(S, O, r, o, d, bounce, max_bounces, record_k, mode, flags, bounce == 0, c)
                           : bounce_once<false>(S, O, r, o, d, bounce, max_bounces, record_k, mode, flags, bounce == 0, c);
        if (alive) {
            R.o[3 * r] = o.x; R.o[3 * r + 1] = o.y; R.o[3 * r + 2] = o.z;
            R.d[3 * r] = d.x; R.d[3 * r + 1] = d.y; R.d[3 * r + 2] = d.z;
        }
    }
    append_live(alive, r, live_out, n_out);
    flush_counters(c, ctr);
}

// -1 / 0 padding of the per-bounce records (coalesced, grid-stride).
__global__ void k_pad_records(Outs O, int64_t total) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        if (O.hit_patch) O.hit_patch[q] = -1;
        if (O.newton_iters) O.newton_iters[q] = 0;
        if (O.hit_xyz) { O.hit_xyz[3 * q] = 0.0; O.hit_xyz[3 * q + 1] = 0.0; O.hit_xyz[3 * q + 2] = 0.0; }
    }
}

__global__ void k_zero_i32(int32_t* p, int n) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = 0;
}

// ---------------------------------------------------------------------------
// Host runtime
// ---------------------------------------------------------------------------
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* dev_upload(const std::vector<T>& v, std::vector<void*>& allocs, int64_t& bytes) {
    T* p = nullptr;
    const size_t sz = std::max<size_t>(1, v.size()) * sizeof(T);
    check(hipMalloc(&p, sz), "hipMalloc(scene)");
    allocs.push_back(p);
    bytes += (int64_t)sz;
    if (!v.empty()) check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(scene)");
    return p;
}

}  // namespace

struct bbrt_scene_s {
    int device = 0;
    int n_cu = 1;
    hipStream_t stream = nullptr;
    DevScene dev{};
    std::vector<void*> allocs;
    int64_t bytes = 0;
    int64_t n_patches = 0;
    // host copies for the query calls
    std::vector<double> pbox, pcent, max_sag;
    std::vector<BvhNode> bvh;
    std::vector<int32_t> leaf;
    double diag = 0.0, max_patch_sag = 0.0;
    // per-trace scratch, grown on demand
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    int32_t* h_count = nullptr;   // pinned
    hipEvent_t ev[2] = { nullptr, nullptr };
    std::vector<hipEvent_t> kev;

    ~bbrt_scene_s() {
        for (void* p : allocs) (void)hipFree(p);
        if (scratch) (void)hipFree(scratch);
        if (h_count) (void)hipHostFree(h_count);
        for (auto e : ev) if (e) (void)hipEventDestroy(e);
        for (auto e : kev) if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int fail(const std::exception& e) {
    g_last_error = e.what();
    return -1;
}

void* scratch_for(bbrt_scene s, size_t need) {
    if (need > s->scratch_bytes) {
        if (s->scratch) check(hipFree(s->scratch), "hipFree(scratch)");
        s->scratch = nullptr;
        s->scratch_bytes = 0;
        check(hipMalloc(&s->scratch, need), "hipMalloc(scratch)");
        s->scratch_bytes = need;
    }
    return s->scratch;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

void trace_impl(bbrt_scene s, const double* d_o, const double* d_d, int64_t n, const bbrt_params* p,
                const bbrt_outputs* out, int64_t* counters, hipStream_t st, double* kernel_ms, bool pad_records) {
    if (!p) throw std::runtime_error("bbrt_trace: params is NULL");
    if (p->max_bounces < 0 || p->record_k < 0) throw std::runtime_error("bbrt_trace: negative max_bounces or record_k");
    if (p->newton_mode != BBRT_NEWTON_REF && p->newton_mode != BBRT_NEWTON_CONVERGED)
        throw std::runtime_error("bbrt_trace: unknown newton_mode");
    if (p->seed_mode != BBRT_SEED_TRIANGLE)
        throw std::runtime_error("bbrt_trace: only the triangle (stage-1) seed is available in this library");
    if (n < 0 || n > (int64_t)INT32_MAX - 1024) throw std::runtime_error("bbrt_trace: ray count out of range");
    const int block = p->block > 0 ? p->block : 256;
    if (block % 64 != 0 || block > 256) throw std::runtime_error("bbrt_trace: block must be a multiple of 64 <= 256");
    if (!st) st = s->stream;
    check(hipSetDevice(s->device), "hipSetDevice");
    int64_t local_ctr[BBRT_NCOUNTERS] = { 0 };
    if (n == 0) {
        if (counters) std::memcpy(counters, local_ctr, sizeof(local_ctr));
        if (kernel_ms) { kernel_ms[0] = 0; kernel_ms[1] = 0; kernel_ms[2] = 0; }
        return;
    }
    // scratch: state o,d (48 B/ray), two live lists, counts, counters
    const size_t off_o = 0;
    const size_t off_d = align256(off_o + (size_t)n * 24);
    const size_t off_la = align256(off_d + (size_t)n * 24);
    const size_t off_lb = align256(off_la + (size_t)n * 4);
    const size_t off_cnt = align256(off_lb + (size_t)n * 4);
    const size_t off_ctr = align256(off_cnt + 256);
    const size_t need = align256(off_ctr + BBRT_NCOUNTERS * 8);
    char* base = static_cast<char*>(scratch_for(s, need));
    RayState R{ reinterpret_cast<double*>(base + off_o), reinterpret_cast<double*>(base + off_d) };
    int32_t* live[2] = { reinterpret_cast<int32_t*>(base + off_la), reinterpret_cast<int32_t*>(base + off_lb) };
    int32_t* cnt = reinterpret_cast<int32_t*>(base + off_cnt);   // [0], [1]: list sizes, [2]: queue head
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(base + off_ctr);
    Outs O{ out ? out->status : nullptr, out ? out->err_kind : nullptr, out ? out->bounces : nullptr,
            out ? out->err_patch : nullptr, out ? out->hit_patch : nullptr, out ? out->hit_xyz : nullptr,
            out ? out->newton_iters : nullptr, out ? out->final_od : nullptr,
            out ? reinterpret_cast<unsigned long long*>(out->hit_count) : nullptr, out ? out->ray_work : nullptr };
    const bool timing = (p->flags & BBRT_FLAG_TIME_KERNELS) != 0;
    const bool wavefront = (p->flags & BBRT_FLAG_WAVEFRONT) != 0;
    if (timing && (int)s->kev.size() < 2 * p->max_bounces + 2) {
        const size_t want = 2 * (size_t)p->max_bounces + 2;
        while (s->kev.size() < want) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            s->kev.push_back(e);
        }
    }
    check(hipEventRecord(s->ev[0], st), "hipEventRecord");
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, cnt, 4);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, reinterpret_cast<int32_t*>(ctr), 2 * BBRT_NCOUNTERS);
    if (pad_records && p->record_k > 0) {
        const int64_t total = n * (int64_t)p->record_k;
        const unsigned pg = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(k_pad_records, dim3(pg), dim3(256), 0, st, O, total);
    }
    const int64_t grid0 = (n + 255) / 256;
    hipLaunchKernelGGL(k_init, dim3((unsigned)grid0), dim3(256), 0, st, d_o, d_d, n, s->dev.bvh, R, O,
                       p->max_bounces, live[0], cnt, ctr);
    check(hipGetLastError(), "k_init launch");
    int launches = 0;
    if (!wavefront) {
        if (p->max_bounces > 0) {
            const int wpe = p->waves_per_eu > 0 ? p->waves_per_eu : kDefaultWavesPerEu;
            decltype(&k_persist<1, false>) kern = nullptr;
            const bool work = O.ray_work != nullptr;
            switch (wpe) {
                case 1: case 2: kern = work ? &k_persist<1, true> : &k_persist<1, false>; break;
                case 3: kern = work ? &k_persist<3, true> : &k_persist<3, false>; break;
                case 4: kern = work ? &k_persist<4, true> : &k_persist<4, false>; break;
                default: throw std::runtime_error("bbrt_trace: waves_per_eu must be 0, 2, 3 or 4");
            }
            int per_cu = 0;
            check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, 0), "occupancy");
            const int64_t resident = (int64_t)std::max(1, per_cu) * s->n_cu;
            const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, resident));
            if (timing) check(hipEventRecord(s->kev[0], st), "hipEventRecord");
            hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, s->dev, R, live[0], cnt, cnt + 2, O,
                               p->max_bounces, p->record_k, p->newton_mode, p->flags, ctr);
            if (timing) check(hipEventRecord(s->kev[1], st), "hipEventRecord");
            check(hipGetLastError(), "k_persist launch");
            launches = 1;
        }
    } else {
        int cur = 0;
        for (int b = 0; b < p->max_bounces; ++b) {
            check(hipMemcpyAsync(s->h_count, cnt + cur, sizeof(int32_t), hipMemcpyDeviceToHost, st), "count D2H");
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            const int32_t n_live = *s->h_count;
            if (n_live <= 0) break;
            hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, cnt + (cur ^ 1), 1);
            const unsigned grid = (unsigned)((n_live + block - 1) / block);
            if (timing) check(hipEventRecord(s->kev[2 * launches], st), "hipEventRecord");
            hipLaunchKernelGGL(k_bounce, dim3(grid), dim3(block), 0, st, s->dev, R, live[cur], cnt + cur,
                               live[cur ^ 1], cnt + (cur ^ 1), O, b, p->max_bounces, p->record_k, p->newton_mode,
                               p->flags, ctr);
            if (timing) check(hipEventRecord(s->kev[2 * launches + 1], st), "hipEventRecord");
            check(hipGetLastError(), "k_bounce launch");
            ++launches;
            cur ^= 1;
        }
    }
    check(hipEventRecord(s->ev[1], st), "hipEventRecord");
    unsigned long long hc[BBRT_NCOUNTERS];
    check(hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st), "counters D2H");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (int q = 0; q < BBRT_NCOUNTERS; ++q) local_ctr[q] = (int64_t)hc[q];
    local_ctr[bbrt::kCtrLaunches] = launches;
    if (counters) std::memcpy(counters, local_ctr, sizeof(local_ctr));
    if (kernel_ms) {
        double ksum = 0.0;
        if (timing)
            for (int q = 0; q < launches; ++q) {
                float ms = 0.0f;
                check(hipEventElapsedTime(&ms, s->kev[2 * q], s->kev[2 * q + 1]), "hipEventElapsedTime");
                ksum += ms;
            }
        float total = 0.0f;
        check(hipEventElapsedTime(&total, s->ev[0], s->ev[1]), "hipEventElapsedTime");
        kernel_ms[0] = ksum; kernel_ms[1] = launches; kernel_ms[2] = total;
    }
}

}  // namespace

extern "C" {

const char* bbrt_last_error(void) { return g_last_error.c_str(); }

int bbrt_prepare_patches(const bbrt_patch_set* ps, double* box, double* centroid, double* max_sagitta,
                         double* coarse_xyz, double* cell) {
    if (!ps) { g_last_error = "bbrt_prepare_patches: NULL patch set"; return -1; }
    try {
        bbrt::HostScene H;
        const unsigned hw = std::thread::hardware_concurrency();
        bbrt::build_host_scene(*ps, H, (int)std::min(16u, std::max(1u, hw)));
        if (box) std::memcpy(box, H.pbox.data(), H.pbox.size() * sizeof(double));
        if (centroid) std::memcpy(centroid, H.pcent.data(), H.pcent.size() * sizeof(double));
        if (max_sagitta) for (int64_t k = 0; k < H.n; ++k) max_sagitta[k] = H.meta[k].max_sag;
        if (coarse_xyz) std::memcpy(coarse_xyz, H.cxyz.data(), H.cxyz.size() * sizeof(double));
        if (cell) std::memcpy(cell, H.cell.data(), H.cell.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int64_t bbrt_build_bvh(const double* box, const double* centroid, int64_t n, int32_t leaf_threshold,
                       double* node_box, int32_t* right, int32_t* start, int32_t* count, uint8_t* axis,
                       int32_t* leaf_objects) {
    if (!box || !centroid || n <= 0 || leaf_threshold < 1) {
        g_last_error = "bbrt_build_bvh: bad arguments";
        return -1;
    }
    try {
        std::vector<double> b(box, box + 6 * n), c(centroid, centroid + 3 * n);
        std::vector<BvhNode> nodes;
        std::vector<int32_t> leaf;
        bbrt::build_bvh(b, c, leaf_threshold, nodes, leaf);
        for (size_t k = 0; k < nodes.size(); ++k) {
            const BvhNode& nd = nodes[k];
            if (node_box) for (int a = 0; a < 3; ++a) { node_box[6 * k + a] = nd.lo[a]; node_box[6 * k + 3 + a] = nd.hi[a]; }
            if (right) right[k] = nd.right;
            if (start) start[k] = nd.start;
            if (count) count[k] = nd.count;
            if (axis) axis[k] = (uint8_t)nd.axis;
        }
        if (leaf_objects) std::memcpy(leaf_objects, leaf.data(), leaf.size() * sizeof(int32_t));
        return (int64_t)nodes.size();
    } catch (const std::exception& e) {
        fail(e);
        return -1;
    }
}

int bbrt_scene_create(const bbrt_patch_set* ps, bbrt_scene* out) {
    if (!ps || !out) { g_last_error = "bbrt_scene_create: NULL argument"; return -1; }
    *out = nullptr;
    bbrt_scene s = nullptr;
    try {
        int ndev = 0;
        check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (ndev <= 0) throw std::runtime_error("bbrt_scene_create: no HIP device");
        bbrt::HostScene H;
        const unsigned hw = std::thread::hardware_concurrency();
        bbrt::build_host_scene(*ps, H, (int)std::min(16u, std::max(1u, hw)));
        s = new bbrt_scene_s();
        check(hipGetDevice(&s->device), "hipGetDevice");
        check(hipDeviceGetAttribute(&s->n_cu, hipDeviceAttributeMultiprocessorCount, s->device), "CU count");
        for (const BvhNode& nd : H.bvh)
            if (nd.right >= (1 << 28) || nd.start >= (1 << 27) || nd.count > 7)
                throw std::runtime_error("bbrt_scene_create: BVH exceeds the packed node limits");
        check(hipStreamCreate(&s->stream), "hipStreamCreate");
        check(hipEventCreate(&s->ev[0]), "hipEventCreate");
        check(hipEventCreate(&s->ev[1]), "hipEventCreate");
        check(hipHostMalloc(reinterpret_cast<void**>(&s->h_count), sizeof(int32_t) * 64, hipHostMallocDefault),
              "hipHostMalloc");
        DevScene& D = s->dev;
        D.meta = dev_upload(H.meta, s->allocs, s->bytes);
        D.nodes = dev_upload(H.nodes, s->allocs, s->bytes);
        D.fields = dev_upload(H.fields, s->allocs, s->bytes);
        D.cxyz = dev_upload(H.cxyz, s->allocs, s->bytes);
        D.cell = dev_upload(H.cell, s->allocs, s->bytes);
        D.box_cell = dev_upload(H.box_cell, s->allocs, s->bytes);
        D.box_tile = dev_upload(H.box_tile, s->allocs, s->bytes);
        D.box_block = dev_upload(H.box_block, s->allocs, s->bytes);
        D.box_super = dev_upload(H.box_super, s->allocs, s->bytes);
        D.box_patch = dev_upload(H.box_patch, s->allocs, s->bytes);
        D.bvh = dev_upload(H.bvh, s->allocs, s->bytes);
        D.leaf = dev_upload(H.leaf, s->allocs, s->bytes);
        D.diag = H.diag;
        D.cull_limit = 8.0 * H.scale;
        s->n_patches = H.n;
        s->pbox = std::move(H.pbox);
        s->pcent = std::move(H.pcent);
        s->max_sag.resize(H.n);
        for (int64_t k = 0; k < H.n; ++k) s->max_sag[k] = H.meta[k].max_sag;
        s->bvh = std::move(H.bvh);
        s->leaf = std::move(H.leaf);
        s->diag = H.diag;
        s->max_patch_sag = H.max_patch_sag;
        *out = s;
        return 0;
    } catch (const std::exception& e) {
        delete s;
        return fail(e);
    }
}

void bbrt_scene_destroy(bbrt_scene s) { delete s; }

int bbrt_scene_info(bbrt_scene s, double* diag, double* max_patch_sag, int64_t* n_nodes, int64_t* n_patches,
                    int64_t* device_bytes) {
    if (!s) { g_last_error = "bbrt_scene_info: NULL scene"; return -1; }
    if (diag) *diag = s->diag;
    if (max_patch_sag) *max_patch_sag = s->max_patch_sag;
    if (n_nodes) *n_nodes = (int64_t)s->bvh.size();
    if (n_patches) *n_patches = s->n_patches;
    if (device_bytes) *device_bytes = s->bytes;
    return 0;
}

int bbrt_scene_patch_info(bbrt_scene s, double* box, double* centroid, double* max_sagitta) {
    if (!s) { g_last_error = "bbrt_scene_patch_info: NULL scene"; return -1; }
    if (box) std::memcpy(box, s->pbox.data(), s->pbox.size() * sizeof(double));
    if (centroid) std::memcpy(centroid, s->pcent.data(), s->pcent.size() * sizeof(double));
    if (max_sagitta) std::memcpy(max_sagitta, s->max_sag.data(), s->max_sag.size() * sizeof(double));
    return 0;
}

int bbrt_scene_bvh(bbrt_scene s, double* node_box, int32_t* right, int32_t* start, int32_t* count, uint8_t* axis,
                   int32_t* leaf_objects) {
    if (!s) { g_last_error = "bbrt_scene_bvh: NULL scene"; return -1; }
    for (size_t k = 0; k < s->bvh.size(); ++k) {
        const BvhNode& nd = s->bvh[k];
        if (node_box) for (int a = 0; a < 3; ++a) { node_box[6 * k + a] = nd.lo[a]; node_box[6 * k + 3 + a] = nd.hi[a]; }
        if (right) right[k] = nd.right;
        if (start) start[k] = nd.start;
        if (count) count[k] = nd.count;
        if (axis) axis[k] = (uint8_t)nd.axis;
    }
    if (leaf_objects) std::memcpy(leaf_objects, s->leaf.data(), s->leaf.size() * sizeof(int32_t));
    return 0;
}

int bbrt_scene_coarse(bbrt_scene s, int64_t k, double* xyz, double* sag, double* cnorm) {
    if (!s) { g_last_error = "bbrt_scene_coarse: NULL scene"; return -1; }
    if (k < 0 || k >= s->n_patches) { g_last_error = "bbrt_scene_coarse: patch index out of range"; return -1; }
    try {
        check(hipSetDevice(s->device), "hipSetDevice");
        const size_t nv = (size_t)bbrt::kCoarseN * bbrt::kCoarseN * 3;
        const size_t nc = (size_t)bbrt::kCells * bbrt::kCells;
        if (xyz) check(hipMemcpy(xyz, s->dev.cxyz + (size_t)k * nv, nv * sizeof(double), hipMemcpyDeviceToHost), "copy");
        if (sag || cnorm) {
            std::vector<double> c(nc * 4);
            check(hipMemcpy(c.data(), s->dev.cell + (size_t)k * nc * 4, nc * 4 * sizeof(double), hipMemcpyDeviceToHost),
                  "copy");
            for (size_t q = 0; q < nc; ++q) {
                if (sag) sag[q] = c[4 * q];
                if (cnorm) { cnorm[3 * q] = c[4 * q + 1]; cnorm[3 * q + 1] = c[4 * q + 2]; cnorm[3 * q + 2] = c[4 * q + 3]; }
            }
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_trace_device(bbrt_scene s, const double* origins, const double* dirs, int64_t n, const bbrt_params* p,
                      const bbrt_outputs* out, int64_t* counters, void* stream, double* kernel_ms) {
    if (!s) { g_last_error = "bbrt_trace_device: NULL scene"; return -1; }
    try {
        trace_impl(s, origins, dirs, n, p, out, counters, static_cast<hipStream_t>(stream), kernel_ms, false);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_trace(bbrt_scene s, const double* origins, const double* dirs, int64_t n, const bbrt_params* p,
               const bbrt_outputs* out, int64_t* counters) {
    if (!s) { g_last_error = "bbrt_trace: NULL scene"; return -1; }
    if (!p) { g_last_error = "bbrt_trace: params is NULL"; return -1; }
    if (n > 0 && (!origins || !dirs)) { g_last_error = "bbrt_trace: NULL origins or dirs"; return -1; }
    std::vector<void*> tmp;
    auto cleanup = [&]() { for (void* q : tmp) (void)hipFree(q); };
    try {
        check(hipSetDevice(s->device), "hipSetDevice");
        const int64_t K = p->record_k;
        auto dalloc = [&](size_t bytes) -> void* {
            void* q = nullptr;
            check(hipMalloc(&q, std::max<size_t>(bytes, 8)), "hipMalloc(trace)");
            tmp.push_back(q);
            return q;
        };
        const size_t nn = (size_t)std::max<int64_t>(n, 0);
        double* d_o = static_cast<double*>(dalloc(nn * 24));
        double* d_d = static_cast<double*>(dalloc(nn * 24));
        if (nn) {
            check(hipMemcpy(d_o, origins, nn * 24, hipMemcpyHostToDevice), "H2D origins");
            check(hipMemcpy(d_d, dirs, nn * 24, hipMemcpyHostToDevice), "H2D dirs");
        }
        bbrt_outputs dout{};
        bbrt_outputs hout = out ? *out : bbrt_outputs{};
        if (hout.status) dout.status = static_cast<int8_t*>(dalloc(nn));
        if (hout.err_kind) dout.err_kind = static_cast<int8_t*>(dalloc(nn));
        if (hout.bounces) dout.bounces = static_cast<int32_t*>(dalloc(nn * 4));
        if (hout.err_patch) dout.err_patch = static_cast<int32_t*>(dalloc(nn * 4));
        if (hout.hit_patch) dout.hit_patch = static_cast<int32_t*>(dalloc(nn * K * 4));
        if (hout.hit_xyz) dout.hit_xyz = static_cast<double*>(dalloc(nn * K * 24));
        if (hout.newton_iters) dout.newton_iters = static_cast<uint8_t*>(dalloc(nn * K));
        if (hout.final_od) dout.final_od = static_cast<double*>(dalloc(nn * 48));
        if (hout.ray_work) dout.ray_work = static_cast<uint32_t*>(dalloc(nn * 16));
        if (hout.hit_count) {
            dout.hit_count = static_cast<int64_t*>(dalloc((size_t)s->n_patches * 8));
            check(hipMemcpy(dout.hit_count, hout.hit_count, (size_t)s->n_patches * 8, hipMemcpyHostToDevice), "H2D hit_count");
        }
        trace_impl(s, d_o, d_d, n, p, &dout, counters, s->stream, nullptr, true);
        auto back = [&](void* h, const void* d, size_t bytes) {
            if (h && bytes) check(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost), "D2H outputs");
        };
        back(hout.status, dout.status, nn);
        back(hout.err_kind, dout.err_kind, nn);
        back(hout.bounces, dout.bounces, nn * 4);
        back(hout.err_patch, dout.err_patch, nn * 4);
        back(hout.hit_patch, dout.hit_patch, nn * K * 4);
        back(hout.hit_xyz, dout.hit_xyz, nn * K * 24);
        back(hout.newton_iters, dout.newton_iters, nn * K);
        back(hout.final_od, dout.final_od, nn * 48);
        back(hout.hit_count, dout.hit_count, (size_t)s->n_patches * 8);
        back(hout.ray_work, dout.ray_work, nn * 16);
        cleanup();
        return 0;
    } catch (const std::exception& e) {
        cleanup();
        return fail(e);
    }
}

}  // extern "C"
