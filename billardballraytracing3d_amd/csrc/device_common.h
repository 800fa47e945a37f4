// device_common.h -- device-side types and functions of the bounce loop, shared by
// the bounce kernels (bbrt_hip.hip) and the learned-seed kernels (bbrt_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "scene_host.h"   // also brings in include/bbrt.h

using bbrt::BvhNode;
using bbrt::PatchMeta;

namespace {

constexpr int kStackCap = 48;
constexpr int kStack4Cap = 64;                // 4-wide traversal stack (3 entries per level; depth <= 21 levels)
constexpr double kPruneSlack = 1.000001;      // FP32 tmin prune: relative slack over the FP64 best_t
constexpr int kStage1Gather = 24;             // leaf lanes a wave gathers before its stage-1 walk: one trace of the
                                              // Sinai scene 1.45 -> 1.33 ms (four in flight: unchanged; 48 gains
                                              // 4 % more alone but loses 2 % in flight), 256 compact sub-scenes
                                              // 5.75 -> 5.51 ms
constexpr int kDefaultWavesPerEu = 2;         // persistent-kernel occupancy target: 2 waves/EU keep the stored-scene
                                              // kernel's spills small and measured 6-7 % faster than 3 on MI355X
constexpr int kDefaultWavesPerEuCompact = 2;  // compact scenes: the on-the-fly cell evaluation wants the registers
inline int default_waves_per_eu(bool compact) { return compact ? kDefaultWavesPerEuCompact : kDefaultWavesPerEu; }
// threads per block of the persistent kernel: 64 for stored scenes with the triangle seed (as fast as 256
// one trace at a time; with traces in flight the next trace's waves start as soon as one wave's slot frees),
// 256 for the learned seed (21 KB of LDS per block) and for compact scenes
inline int default_block(bool learned, bool compact) { return (learned || compact) ? 256 : 64; }
constexpr int kDefaultNewtonBatch = 64;       // parked lanes (stage-1 hits) that trigger a batched Newton (64: when no lane can go on)
constexpr int kDefaultNewtonBatchCompact = 32;  // compact scenes: 32 (256 sub-scenes: 5.04-5.12 vs 5.53-5.60 ms at 64;
                                                // the Sinai scene measures 32 and 64 alike)
inline int default_newton_batch(bool compact) { return compact ? kDefaultNewtonBatchCompact : kDefaultNewtonBatch; }
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
    const bbrt::Bvh4Node* bvh4;      // per binary node: its 4-wide node (filled at even depth)
    const int32_t* leaf;
    const double* leaf_box;          // [leaf position][6]: FP64 box of the leaf whose objects start there
    const float* leaf_pbox;          // [leaf position][8]: culling box of the patch at that leaf position
    const bbrt::SceneInfo* scenes;   // per sub-scene: BVH root, first patch, diagonal, culling limit
    const int32_t* ray_scene;        // per ray: its sub-scene (NULL: every ray in sub-scene 0)
    // compact scenes: no cxyz / cell / box_cell; cells are evaluated from the node-set tables
    const double* ff_rows;           // [set][kCoarseN][kRowStride]
    const double* mid_basis;         // [set][kCells][kMaxNodes1D]
    const uint8_t* qbox_cell;        // [n][kQCellPatchBytes]
    int compact;
};

__device__ __forceinline__ int ray_scene_of(const DevScene& S, int64_t r) { return S.ray_scene ? S.ray_scene[r] : 0; }

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
    uint32_t* seed_stats;
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
    uint32_t seed_learned, seed_fallback, mlp_evals, mlp_lanes;   // learned-seed mode only
    uint32_t newton_iters32;                                      // BBRT_NEWTON_CONVERGED_FP32: FP32 iterations
    uint32_t seed_miss;                                           // BBRT_FLAG_SEED_FIRST: misses without stage 1
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

// FP32 ray test of the four slots of a 4-wide node: the mask of slots whose
// (widened) box the ray enters at t >= 0, and their entry t.  Widening and
// outward rounding make it a superset of the FP64 test of the slots' binary
// nodes for rays with |origin| within the culling limit.
__device__ __forceinline__ uint32_t slots32(const bbrt::Bvh4Node* __restrict__ N, const Ray32& r, float tmin_out[4]) {
    const float4* p = reinterpret_cast<const float4*>(N);
    const float4 lx = p[0], ly = p[1], lz = p[2], hx = p[3], hy = p[4], hz = p[5];
    const float lo[3][4] = { { lx.x, lx.y, lx.z, lx.w }, { ly.x, ly.y, ly.z, ly.w }, { lz.x, lz.y, lz.z, lz.w } };
    const float hi[3][4] = { { hx.x, hx.y, hx.z, hx.w }, { hy.x, hy.y, hy.z, hy.w }, { hz.x, hz.y, hz.z, hz.w } };
    const float oo[3] = { r.ox, r.oy, r.oz }, ii[3] = { r.ix, r.iy, r.iz };
    uint32_t m = 0;
#pragma unroll
    for (int sl = 0; sl < 4; ++sl) {
        float tmin = -INFINITY, tmax = INFINITY;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float t1 = (lo[a][sl] - oo[a]) * ii[a], t2 = (hi[a][sl] - oo[a]) * ii[a];
            if (ii[a] < 0.0f) { const float t = t1; t1 = t2; t2 = t; }
            tmin = (tmin < t1) ? t1 : tmin;
            tmax = (t2 < tmax) ? t2 : tmax;
        }
        tmin_out[sl] = tmin;
        if (tmax >= tmin && tmax >= 0.0f) m |= 1u << sl;
    }
    return m;
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

// ---------------------------------------------------------------------------
// Compact scenes: a coarse cell evaluated from the nodal fields, bit-identical
// to the host build (scene_build.cpp compute_grid / compute_cells / build_pyramid).
// ---------------------------------------------------------------------------
struct CellData {
    V3 p00, p10, p01, p11;   // corners (i,j), (i+1,j), (i,j+1), (i+1,j+1)
    double s;                // sagitta
    V3 n;                    // unit normal at the cell midpoint
};

// Coarse vertex (i, j): (sum_m (sum_q ru[i][q] f[q][m]) rv[j][m]) * (su[i] sv[j]) per coordinate,
// first-form interpolation (PatchInterpolation.hpp:54-159) with the inner index summed upwards.
// The four corners of a cell are evaluated together, one coordinate at a
// time: the inner sums over q for rows i and i+1 are shared by the corners of
// equal i.  Each corner's operations and their order are those of the
// single-vertex form, so the values are bit-identical.
// N > 0: N x N nodes known at compile time; the loops unroll, so one
// coordinate's nodal values are loaded together (one memory round trip).
// N == 0: any node counts, plain loops.
template <int N>
__device__ __forceinline__ void corner_coord(const double* __restrict__ f, const double* __restrict__ ru0,
                                             const double* __restrict__ ru1, const double* __restrict__ rv0,
                                             const double* __restrict__ rv1, int im_, int jm_, double a[4]) {
    const int im = N > 0 ? N : im_, jm = N > 0 ? N : jm_;
    constexpr int L = N > 0 ? N : 1;
    double a00 = 0.0, a10 = 0.0, a01 = 0.0, a11 = 0.0;
    if constexpr (N > 0) {
#pragma unroll
        for (int m = 0; m < L; ++m) {
            double c0 = 0.0, c1 = 0.0;
#pragma unroll
            for (int q = 0; q < L; ++q) {
                const double fv = f[q * L + m];
                c0 += ru0[q] * fv;
                c1 += ru1[q] * fv;
            }
            const double v0 = rv0[m], v1 = rv1[m];
            a00 += c0 * v0; a10 += c1 * v0; a01 += c0 * v1; a11 += c1 * v1;
        }
    } else {
        for (int m = 0; m < jm; ++m) {
            double c0 = 0.0, c1 = 0.0;
            for (int q = 0; q < im; ++q) {
                const double fv = f[q * jm + m];
                c0 += ru0[q] * fv;
                c1 += ru1[q] * fv;
            }
            const double v0 = rv0[m], v1 = rv1[m];
            a00 += c0 * v0; a10 += c1 * v0; a01 += c0 * v1; a11 += c1 * v1;
        }
    }
    a[0] = a00; a[1] = a10; a[2] = a01; a[3] = a11;
}

// field value at the cell midpoint: sum_i lu[i] (sum_j lv[j] f[i][j]), rows with lu[i] == 0 skipped
// (scene_build.cpp compute_cells does the same).
template <int N>
__device__ __forceinline__ double field_tab(const double* lu, const double* lv, const double* f, int im_, int jm_) {
    const int im = N > 0 ? N : im_, jm = N > 0 ? N : jm_;
    constexpr int L = N > 0 ? N : 1;
    double sum = 0.0;
    if constexpr (N > 0) {
#pragma unroll
        for (int i = 0; i < L; ++i) {
            const double li = lu[i];
            double row = 0.0;
#pragma unroll
            for (int j = 0; j < L; ++j) row += lv[j] * f[i * L + j];
            if (li != 0.0) sum += li * row;
        }
    } else {
        for (int i = 0; i < im; ++i) {
            const double li = lu[i];
            if (li != 0.0) {
                double row = 0.0;
                for (int j = 0; j < jm; ++j) row += lv[j] * f[i * jm + j];
                sum += li * row;
            }
        }
    }
    return sum;
}

// Corners, sagitta and normal of cell (i, j) (SurfacePatch.hpp:416-492).
// The loops over coordinates and fields stay rolled (values rotate through
// registers instead of being indexed), so at most one field's nodal values are
// live at a time; the fully unrolled form spills in the persistent kernel.
template <int N>
__device__ CellData compact_cell(const DevScene& S, int64_t k, int i, int j) {
    const int32_t ns = S.meta[k].nodeset;
    const int64_t foff = S.meta[k].field_off;
    const int im = N > 0 ? N : S.meta[k].imax, jm = N > 0 ? N : S.meta[k].jmax, nn = im * jm;
    const double* ru0 = S.ff_rows + ((size_t)(ns >> 16) * bbrt::kCoarseN + i) * bbrt::kRowStride;
    const double* ru1 = ru0 + bbrt::kRowStride;
    const double* rv0 = S.ff_rows + ((size_t)(ns & 0xFFFF) * bbrt::kCoarseN + j) * bbrt::kRowStride;
    const double* rv1 = rv0 + bbrt::kRowStride;
    const double* F = S.fields + foff;
    const double* lu = S.mid_basis + ((size_t)(ns >> 16) * bbrt::kCells + i) * bbrt::kMaxNodes1D;
    const double* lv = S.mid_basis + ((size_t)(ns & 0xFFFF) * bbrt::kCells + j) * bbrt::kMaxNodes1D;
    CellData cd;
    // position fields: the corners and the midpoint value from the same nodal values
    double x00 = 0, y00 = 0, z00 = 0, x10 = 0, y10 = 0, z10 = 0, x01 = 0, y01 = 0, z01 = 0, x11 = 0, y11 = 0, z11 = 0;
    double t0 = 0, t1 = 0, t2 = 0;
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        double a[4];
        corner_coord<N>(F + c * nn, ru0, ru1, rv0, rv1, im, jm, a);
        const double tv = field_tab<N>(lu, lv, F + c * nn, im, jm);
        x00 = y00; y00 = z00; z00 = a[0];
        x10 = y10; y10 = z10; z10 = a[1];
        x01 = y01; y01 = z01; z01 = a[2];
        x11 = y11; y11 = z11; z11 = a[3];
        t0 = t1; t1 = t2; t2 = tv;
    }
    {
        const double u0 = ru0[bbrt::kMaxNodes1D], u1 = ru1[bbrt::kMaxNodes1D];
        const double v0 = rv0[bbrt::kMaxNodes1D], v1 = rv1[bbrt::kMaxNodes1D];
        const double s00 = u0 * v0, s10 = u1 * v0, s01 = u0 * v1, s11 = u1 * v1;
        cd.p00 = { x00 * s00, y00 * s00, z00 * s00 };
        cd.p10 = { x10 * s10, y10 * s10, z10 * s10 };
        cd.p01 = { x01 * s01, y01 * s01, z01 * s01 };
        cd.p11 = { x11 * s11, y11 * s11, z11 * s11 };
    }
    // the normal fields (9-11) at the midpoint
    double n0 = 0, n1 = 0, n2 = 0;
#pragma unroll 1
    for (int e = 9; e < 12; ++e) {
        const double v = field_tab<N>(lu, lv, F + e * nn, im, jm);
        n0 = n1; n1 = n2; n2 = v;
    }
    const V3 tm = { t0, t1, t2 };
    cd.n = unit3({ n0, n1, n2 });
    const V3 bm = { 0.25 * (cd.p00.x + cd.p10.x + cd.p01.x + cd.p11.x), 0.25 * (cd.p00.y + cd.p10.y + cd.p01.y + cd.p11.y),
                    0.25 * (cd.p00.z + cd.p10.z + cd.p01.z + cd.p11.z) };
    const V3 dd = sub3(tm, bm);
    cd.s = sqrt(dd.x * dd.x + dd.y * dd.y + dd.z * dd.z);
    return cd;
}

__device__ __forceinline__ CellData compact_cell_any(const DevScene& S, int64_t k, int i, int j) {
    const PatchMeta M = S.meta[k];
    if (M.imax == 3 && M.jmax == 3) return compact_cell<3>(S, k, i, j);
    if (M.imax == 5 && M.jmax == 5) return compact_cell<5>(S, k, i, j);
    return compact_cell<0>(S, k, i, j);
}

__device__ __forceinline__ float round_down_f(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}
__device__ __forceinline__ float round_up_f(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

// The cell's FP32 culling box exactly as build_pyramid makes it.
__device__ __forceinline__ void compact_cell_box_f(const CellData& cd, double widen, float b[6]) {
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    const double signs[3] = { 0.0, 1.0, -1.0 };
    const V3* corner[4] = { &cd.p00, &cd.p10, &cd.p01, &cd.p11 };
    for (int si = 0; si < 3; ++si) {
        const double dd = signs[si] * cd.s;
        for (int q = 0; q < 4; ++q) {
            const double x = corner[q]->x + cd.n.x * dd, y = corner[q]->y + cd.n.y * dd, z = corner[q]->z + cd.n.z * dd;
            lo[0] = (x < lo[0]) ? x : lo[0]; hi[0] = (hi[0] < x) ? x : hi[0];
            lo[1] = (y < lo[1]) ? y : lo[1]; hi[1] = (hi[1] < y) ? y : hi[1];
            lo[2] = (z < lo[2]) ? z : lo[2]; hi[2] = (hi[2] < z) ? z : hi[2];
        }
    }
    b[0] = round_down_f(lo[0] - widen); b[1] = round_down_f(lo[1] - widen); b[2] = round_down_f(lo[2] - widen);
    b[3] = round_up_f(hi[0] + widen); b[4] = round_up_f(hi[1] + widen); b[5] = round_up_f(hi[2] + widen);
}

// ... and the slab test against it.
__device__ __forceinline__ bool compact_cell_box(const CellData& cd, double widen, const Ray32& r) {
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    const double signs[3] = { 0.0, 1.0, -1.0 };
    const V3* corner[4] = { &cd.p00, &cd.p10, &cd.p01, &cd.p11 };
    for (int si = 0; si < 3; ++si) {
        const double dd = signs[si] * cd.s;
        for (int q = 0; q < 4; ++q) {
            const double x = corner[q]->x + cd.n.x * dd, y = corner[q]->y + cd.n.y * dd, z = corner[q]->z + cd.n.z * dd;
            lo[0] = (x < lo[0]) ? x : lo[0]; hi[0] = (hi[0] < x) ? x : hi[0];
            lo[1] = (y < lo[1]) ? y : lo[1]; hi[1] = (hi[1] < y) ? y : hi[1];
            lo[2] = (z < lo[2]) ? z : lo[2]; hi[2] = (hi[2] < z) ? z : hi[2];
        }
    }
    const float b[6] = { round_down_f(lo[0] - widen), round_down_f(lo[1] - widen), round_down_f(lo[2] - widen),
                         round_up_f(hi[0] + widen), round_up_f(hi[1] + widen), round_up_f(hi[2] + widen) };
    return box32(b, r);
}

// The six Moller-Trumbore tests of a cell given its data (SurfacePatch.hpp:274-294).
__device__ __forceinline__ void cell_tests_data(const CellData& cd, int i, int j, const V3& o, const V3& d,
                                                Stage1& best, Counters& c) {
    const int base = (i * bbrt::kCells + j) * 6;
    const double signs[3] = { 0.0, 1.0, -1.0 };
#pragma unroll
    for (int si = 0; si < 3; ++si) {
        const double dd = signs[si] * cd.s;
        const V3 q00 = { cd.p00.x + cd.n.x * dd, cd.p00.y + cd.n.y * dd, cd.p00.z + cd.n.z * dd };
        const V3 q10 = { cd.p10.x + cd.n.x * dd, cd.p10.y + cd.n.y * dd, cd.p10.z + cd.n.z * dd };
        const V3 q01 = { cd.p01.x + cd.n.x * dd, cd.p01.y + cd.n.y * dd, cd.p01.z + cd.n.z * dd };
        const V3 q11 = { cd.p11.x + cd.n.x * dd, cd.p11.y + cd.n.y * dd, cd.p11.z + cd.n.z * dd };
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

// Stage 1 through the culling pyramid (grouped layout: one load per level), stored scenes.
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

// Cells of tile (ti, tj) whose quantized box (a superset of the exact FP32 box) the ray enters.
__device__ __forceinline__ uint32_t qcell_mask(const DevScene& S, int64_t k, const float* tb, int ti, int tj,
                                               const Ray32& r) {
    const uint2* q8 = reinterpret_cast<const uint2*>(S.qbox_cell + (size_t)k * bbrt::kQCellPatchBytes +
                                                     (size_t)(ti * bbrt::kTileN + tj) * bbrt::kQCellBytes);
    uint32_t qw[14];   // the 56 bytes as little-endian words
#pragma unroll
    for (int w = 0; w < 7; ++w) {
        const uint2 v = q8[w];
        qw[2 * w] = v.x; qw[2 * w + 1] = v.y;
    }
    const double tlo[3] = { tb[0], tb[1], tb[2] };
    const double step[3] = { ((double)tb[3] - tlo[0]) / 255.0, ((double)tb[4] - tlo[1]) / 255.0,
                             ((double)tb[5] - tlo[2]) / 255.0 };
    uint32_t mc = 0;
#pragma unroll
    for (int cq = 0; cq < 9; ++cq) {
        double qv[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const int byte = 6 * cq + e;
            qv[e] = (double)((qw[byte >> 2] >> (8 * (byte & 3))) & 0xFFu);
        }
        const float bx[6] = { round_down_f(tlo[0] + qv[0] * step[0]), round_down_f(tlo[1] + qv[1] * step[1]),
                              round_down_f(tlo[2] + qv[2] * step[2]), round_up_f(tlo[0] + qv[3] * step[0]),
                              round_up_f(tlo[1] + qv[4] * step[1]), round_up_f(tlo[2] + qv[5] * step[2]) };
        if (box32(bx, r)) mc |= 1u << cq;
    }
    return mc;
}

// The resumable walk of a compact patch's culling pyramid, one lane's state.
// Culled: the pyramid's super, block and tile boxes select tiles, the
// quantized boxes select the tile's cells.  Not culled: every cell, tile by
// tile.  Tiles come in the stored path's order.
struct CompactWalk {
    uint32_t ms, mb, mt;
    int si, sj, bi, bj, t;
    bool use_cull;
};

__device__ __forceinline__ void compact_walk_init(const DevScene& S, int64_t k, const Ray32& r, bool use_cull,
                                                  CompactWalk& w, Counters& c) {
    w = CompactWalk{ 0u, 0u, 0u, 0, 0, 0, 0, 0, use_cull };
    if (use_cull) {
        w.ms = group_mask<9, bbrt::kG9>(S.box_super + (size_t)k * bbrt::kSuperFloats, r);
        c.cull_tests += 9;
    }
}

// The next tile with a selected cell: true with the tile (ti, tj) and its
// cell mask mc (bit 3 * (i % 3) + j % 3), false when the walk is over.
__device__ __forceinline__ bool compact_walk_next(const DevScene& S, int64_t k, const Ray32& r, CompactWalk& w,
                                                  int& ti, int& tj, uint32_t& mc, Counters& c) {
    const float* gb = S.box_block + (size_t)k * bbrt::kBlockFloats;
    const float* gt = S.box_tile + (size_t)k * bbrt::kTileFloats;
    for (;;) {
        if (w.use_cull) {
            while (w.mt == 0) {
                while (w.mb == 0) {
                    if (w.ms == 0) return false;
                    const int sq = __ffs(w.ms) - 1;
                    w.ms &= w.ms - 1;
                    w.si = sq / 3; w.sj = sq % 3;
                    w.mb = group_mask<16, bbrt::kG16>(gb + sq * bbrt::kG16, r);
                    c.cull_tests += 16;
                }
                const int bq = __ffs(w.mb) - 1;
                w.mb &= w.mb - 1;
                w.bi = bbrt::kBlocksPerSuper * w.si + (bq >> 2); w.bj = bbrt::kBlocksPerSuper * w.sj + (bq & 3);
                w.mt = group_mask<9, bbrt::kG9>(gt + (w.bi * bbrt::kBlockN + w.bj) * bbrt::kG9, r);
                c.cull_tests += 9;
            }
            const int tq = __ffs(w.mt) - 1;
            w.mt &= w.mt - 1;
            ti = 3 * w.bi + tq / 3; tj = 3 * w.bj + tq % 3;
            mc = qcell_mask(S, k, gt + (w.bi * bbrt::kBlockN + w.bj) * bbrt::kG9 + 6 * tq, ti, tj, r);
            c.cull_tests += 9;
        } else {
            if (w.t == bbrt::kTileN * bbrt::kTileN) return false;
            ti = w.t / bbrt::kTileN; tj = w.t % bbrt::kTileN;
            ++w.t;
            mc = 0x1FFu;
        }
        if (mc) return true;
    }
}

// A selected cell: evaluated from the fields, its exact FP32 box rebuilt and
// tested as a stored scene's (culled walks), then Moller-Trumbore.
__device__ __forceinline__ void compact_cell_tests(const DevScene& S, int64_t k, int i, int j, const V3& o,
                                                   const V3& d, const Ray32& r, double widen, bool use_cull,
                                                   Stage1& best, Counters& c) {
    const CellData cd = compact_cell_any(S, k, i, j);
    if (!use_cull || compact_cell_box(cd, widen, r)) cell_tests_data(cd, i, j, o, d, best, c);
}

// Stage 1 of a compact scene by one lane: its walk, each selected cell tested
// as it comes.  Ties between equal t are broken by the cell key, so the result
// does not depend on the order the cells are tested in.
__device__ void stage1_compact(const DevScene& S, int64_t k, const V3& o, const V3& d, const Ray32& r, double widen,
                               bool use_cull, Stage1& best, Counters& c) {
    CompactWalk w;
    compact_walk_init(S, k, r, use_cull, w, c);
    int ti, tj;
    uint32_t mc;
    while (compact_walk_next(S, k, r, w, ti, tj, mc, c)) {
        while (mc) {
            const int cq = __ffs(mc) - 1;
            mc &= mc - 1;
            compact_cell_tests(S, k, 3 * ti + cq / 3, 3 * tj + cq % 3, o, d, r, widen, use_cull, best, c);
        }
    }
}

// k-th (0-based) set bit of m, which has more than k bits set.
__device__ __forceinline__ int nth_bit(unsigned long long m, int k) {
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const int mid = lo + step;
        if ((int)__popcll(m & ((1ull << mid) - 1ull)) <= k) lo = mid;
    }
    return lo;
}

// A lane's stage-1 record as the other lanes of its wave read it while they
// test boxes and cells for it (one LDS record per thread of the block): its
// ray, its patch, the culling widening of its scene, its work counts and its
// best (t, key) so far.
struct Owner {
    double ox, oy, oz, dx, dy, dz;
    double t;
    double widen;
    float r32[6];
    int64_t k;
    uint32_t n_cull, n_mt;
    int32_t key, have;
    int32_t use_cull, pad;
};

// The cells the lanes' walks selected, tested by the whole wave (every lane
// calls this; `have` marks the lanes with a tile).  The wave's list holds the
// lanes' cells in lane order, each lane's in bit order; lane e of a pass takes
// list entry e, tests that cell of its owner's patch for its owner's ray, and
// the owners fold the entries' results into their stage-1 best.  The cells
// tested and the counts are the per-lane walk's, and the best is the minimum
// of (t, key), so the result is the per-lane walk's too; but a cell costs one
// lane of a 64-wide pass instead of a pass of its own.  The owners' rays go
// through LDS (`slots`, this wave's 64 entries) and are read only once the
// cell is evaluated, so they hold no registers during its field loads.
__device__ __forceinline__ void compact_cells_wave(const DevScene& S, bool have, int32_t k, int ti, int tj,
                                                   uint32_t mc, const V3& o, const V3& d, const Ray32& r,
                                                   double widen, bool use_cull, Stage1& best, Counters& c,
                                                   Owner* slots) {
    const int lane = __lane_id();
    const int cnt = have ? __popc(mc) : 0;
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    const int total = __shfl(incl, 63);
    if (total == 0) return;
    const int start = incl - cnt;
    if (have) {
        Owner& me = slots[lane];
        me.ox = o.x; me.oy = o.y; me.oz = o.z; me.dx = d.x; me.dy = d.y; me.dz = d.z;
        me.widen = widen;
        me.r32[0] = r.ox; me.r32[1] = r.oy; me.r32[2] = r.oz; me.r32[3] = r.ix; me.r32[4] = r.iy; me.r32[5] = r.iz;
        me.use_cull = use_cull ? 1 : 0;
    }
    // the wave's LDS writes land before any lane of it reads them (one wave, in-order LDS)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int p0 = 0; p0 < total; p0 += 64) {
        const int e = p0 + lane;
        int own = 0;   // lanes whose cells all come before entry e: its owner
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (__shfl(incl, own + step - 1) <= e) own += step;
        const int32_t kk = __shfl(k, own);
        const int oti = __shfl(ti, own), otj = __shfl(tj, own), ost = __shfl(start, own);
        const uint32_t omc = (uint32_t)__shfl((int)mc, own);
        Stage1 ib{ 0.0, 0, false };
        if (e < total) {
            const int cq = nth_bit(omc, e - ost);
            const int ci = 3 * oti + cq / 3, cj = 3 * otj + cq % 3;
            const CellData cd = compact_cell_any(S, kk, ci, cj);
            // the owner's ray only now: it is not held across the cell's field loads
            __builtin_amdgcn_sched_barrier(0);
            const Owner& R = slots[own];
            const Ray32 orr = { R.r32[0], R.r32[1], R.r32[2], R.r32[3], R.r32[4], R.r32[5] };
            if (!R.use_cull || compact_cell_box(cd, R.widen, orr)) {
                const V3 oo = { R.ox, R.oy, R.oz }, od = { R.dx, R.dy, R.dz };
                cell_tests_data(cd, ci, cj, oo, od, ib, c);
            }
        }
        // each owner's entries of this pass: list entries start .. start + cnt - 1
        for (int u = 0; u < 9; ++u) {
            if (__ballot(u < cnt) == 0) break;   // past every lane's last entry
            const int src = start + u - p0;
            const bool take = u < cnt && src >= 0 && src < 64;
            const int sl = src & 63;
            const double t2 = __shfl(ib.t, sl);
            const int k2 = __shfl(ib.key, sl);
            const bool h2 = __shfl(ib.have ? 1 : 0, sl) != 0;
            if (take && h2 && (!best.have || t2 < best.t || (t2 == best.t && k2 < best.key))) {
                best.have = true; best.t = t2; best.key = k2;
            }
        }
    }
    // the slots are written again by the next call: every lane's reads of this call come first
    __builtin_amdgcn_wave_barrier();
}

struct PHit {
    int status;   // 1 hit, 0 miss, -kind error
    double t;
    V3 p, n;
    int iters;
    double u = 0.0, v = 0.0;   // the converged parameters of a hit
    bool polish_ok = true;     // converged mode: the polish met its tolerance
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

// The FP32 start of BBRT_NEWTON_CONVERGED_FP32: the second-form basis, the field
// sum and Cramer's rule of the FP64 path, evaluated in FP32 from the FP64 nodal
// data.  No result of this phase is kept unless the FP64 iterations after it
// confirm it (see newton()).
template <int N>
__device__ __forceinline__ void basis32(const double* __restrict__ nodes, const double* __restrict__ w, int n, float x,
                                        float* b) {
    constexpr int M = N > 0 ? N : bbrt::kMaxNodes1D;
    const int cnt = N > 0 ? N : n;
    int hot = -1;
#pragma unroll
    for (int i = 0; i < M; ++i)
        if (i < cnt && hot < 0 && x == (float)nodes[i]) hot = i;
    if (hot >= 0) {
#pragma unroll
        for (int i = 0; i < M; ++i) b[i] = (i == hot) ? 1.0f : 0.0f;
        return;
    }
    float denom = 0.0f;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        if (i < cnt) {
            const float li = (float)w[i] / (x - (float)nodes[i]);
            b[i] = li;
            denom += li;
        } else {
            b[i] = 0.0f;
        }
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
        if (i < cnt) b[i] /= denom;
}

template <int N>
__device__ __forceinline__ float field32(const float* lu, const float* lv, const double* __restrict__ f, int im, int jm) {
    constexpr int M = N > 0 ? N : bbrt::kMaxNodes1D;
    const int ci = N > 0 ? N : im, cj = N > 0 ? N : jm;
    if constexpr (N > 0) __builtin_amdgcn_sched_barrier(0);
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        if (i < ci) {
            float row = 0.0f;
#pragma unroll
            for (int j = 0; j < M; ++j)
                if (j < cj) row += lv[j] * (float)f[i * cj + j];
            sum += lu[i] * row;
        }
    }
    return sum;
}

struct F3 { float x, y, z; };

// Cramer's rule in FP32; false when the determinant is zero or not finite.
__device__ __forceinline__ bool solve3_32(const float J[3][3], const F3& b, float x[3]) {
    const float a11 = J[0][0], a12 = J[0][1], a13 = J[0][2];
    const float a21 = J[1][0], a22 = J[1][1], a23 = J[1][2];
    const float a31 = J[2][0], a32 = J[2][1], a33 = J[2][2];
    const float det = a11 * (a22 * a33 - a23 * a32) - a12 * (a21 * a33 - a23 * a31) + a13 * (a21 * a32 - a22 * a31);
    if (!(fabsf(det) > 0.0f) || !isfinite(det)) return false;
    const float det0 = b.x * (a22 * a33 - a23 * a32) - a12 * (b.y * a33 - a23 * b.z) + a13 * (b.y * a32 - a22 * b.z);
    const float det1 = a11 * (b.y * a33 - a23 * b.z) - b.x * (a21 * a33 - a23 * a31) + a13 * (a21 * b.z - b.y * a31);
    const float det2 = a11 * (a22 * b.z - b.y * a32) - a12 * (a21 * b.z - b.y * a31) + b.x * (a21 * a32 - a22 * a31);
    x[0] = det0 / det; x[1] = det1 / det; x[2] = det2 / det;
    return true;
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

// Newton on (u, v, t), SurfacePatch.hpp:593-660; the converged modes add the polish.  F32: the kernel carries
// the FP32 start of BBRT_NEWTON_CONVERGED_FP32 (a separate instantiation, so the FP64-only kernels keep their
// registers), for the 3x3 and 5x5 node sets; other node counts run FP64 throughout.
// PROBE: the network-first form's Newton from the network seed -- at most kProbeIters iterations, and it gives
// up (a no-convergence result) as soon as the iterate is farther than kProbeBox from the patch in u or v, so a
// candidate the ray misses costs a few iterations before stage 1 decides it.
constexpr int kProbeIters = 6;
constexpr double kProbeBox = 3.0;
template <int N, bool F32 = false, bool PROBE = false>
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
    if constexpr (F32 && N > 0) if (mode == BBRT_NEWTON_CONVERGED_FP32) {
        // FP32 iterations from the seed until the reference's step criterion holds in FP32; then the FP64
        // loop below continues from there (its first step is then small: it makes the reference's stop,
        // determinant and domain decisions in FP64, and the polish follows).  If the FP32 phase does not get
        // there (no convergence, a zero or non-finite determinant or step), Newton starts over in FP64 from
        // the seed, exactly as in BBRT_NEWTON_CONVERGED.
        float fu = (float)u, fv = (float)v, ft = (float)t;
        const F3 fo = { (float)o.x, (float)o.y, (float)o.z }, fd = { (float)d.x, (float)d.y, (float)d.z };
        float flu[L], flv[L];
        bool ok32 = false;
        int it32 = 0;
        while (it32 < kNewtonMaxIters) {
            basis32<N>(un, uw, im, fu, flu);
            basis32<N>(vn, vw, jm, fv, flv);
            const F3 X = { field32<N>(flu, flv, F, im, jm), field32<N>(flu, flv, F + nn, im, jm),
                           field32<N>(flu, flv, F + 2 * nn, im, jm) };
            const F3 Xu = { field32<N>(flu, flv, F + 3 * nn, im, jm), field32<N>(flu, flv, F + 4 * nn, im, jm),
                            field32<N>(flu, flv, F + 5 * nn, im, jm) };
            const F3 Xv = { field32<N>(flu, flv, F + 6 * nn, im, jm), field32<N>(flu, flv, F + 7 * nn, im, jm),
                            field32<N>(flu, flv, F + 8 * nn, im, jm) };
            const F3 Fv = { X.x - (fo.x + ft * fd.x), X.y - (fo.y + ft * fd.y), X.z - (fo.z + ft * fd.z) };
            const float J[3][3] = { { Xu.x, Xv.x, -fd.x }, { Xu.y, Xv.y, -fd.y }, { Xu.z, Xv.z, -fd.z } };
            float dl[3];
            if (!solve3_32(J, Fv, dl)) break;
            fu -= dl[0]; fv -= dl[1]; ft -= dl[2];
            ++it32;
            const float sn = sqrtf(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
            if (!isfinite(sn)) break;
            if (sn < (float)kNewtonTol) { ok32 = true; break; }
        }
        c.newton_iters32 += it32;
        if (ok32) { u = fu; v = fv; t = ft; }
    }
    for (int it = 0; it < (PROBE ? kProbeIters : kNewtonMaxIters); ++it) {
        double delta[3];
        if (!step(delta)) {
            c.newton_iters += iters; c.newton_flops += iters * I_it;
            r.status = -bbrt::kErrSingular; r.iters = iters; return r;
        }
        u -= delta[0]; v -= delta[1]; t -= delta[2];
        ++iters;
        step_norm = sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
        if (step_norm < kNewtonTol) break;
        if constexpr (PROBE)
            if (!(fabs(u) <= kProbeBox && fabs(v) <= kProbeBox)) break;
    }
    if (step_norm >= kNewtonTol) {
        c.newton_iters += iters; c.newton_flops += iters * I_it;
        r.status = -bbrt::kErrNoConv; r.iters = iters; return r;
    }
    if (u < -1.0 - kDomainTol || u > 1.0 + kDomainTol || v < -1.0 - kDomainTol || v > 1.0 + kDomainTol) {
        c.newton_iters += iters; c.newton_flops += iters * I_it;
        r.status = -bbrt::kErrDomain; r.iters = iters; r.u = u; r.v = v; return r;
    }
    if (mode != BBRT_NEWTON_REF) {
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
        r.polish_ok = ok;
    }
    c.newton_iters += iters; c.newton_flops += iters * I_it + F_fin;
    basis<N>(un, uw, im, u, lu);
    basis<N>(vn, vw, jm, v, lv);
    r.status = 1; r.t = t; r.iters = iters; r.u = u; r.v = v;
    r.p = { field<N>(lu, lv, F, im, jm), field<N>(lu, lv, F + nn, im, jm), field<N>(lu, lv, F + 2 * nn, im, jm) };
    r.n = unit3({ field<N>(lu, lv, F + 9 * nn, im, jm), field<N>(lu, lv, F + 10 * nn, im, jm),
                  field<N>(lu, lv, F + 11 * nn, im, jm) });
    return r;
}

// Stage 1 of SurfacePatch::intersect (SurfacePatch.hpp:272-298) and the
// triangle-centroid seed (triangle_uv_centroid, :499-510; coarse_.u[i] =
// -1 + 2i/99, :421).  have = false when no sub-triangle is hit (a miss).
struct Seed {
    double u, v, t;
    bool have;
};

__device__ __forceinline__ Seed seed_of(const Stage1& best);

template <int CM>
__device__ __forceinline__ Seed stage1_seed(const DevScene& S, int64_t k, const V3& o, const V3& d, const Ray32& r32,
                                            bool cull, double widen, Counters& c) {
    const double* G = S.cxyz + (size_t)k * bbrt::kCoarseN * bbrt::kCoarseN * 3;
    const double* C = S.cell + (size_t)k * bbrt::kCells * bbrt::kCells * 4;
    Stage1 best{ 0.0, 0, false };
    c.candidates++;
    const bool use_cull = cull && S.meta[k].cull;
    if (CM == 1 || (CM == 2 && S.compact)) stage1_compact(S, k, o, d, r32, widen, use_cull, best, c);
    else if (use_cull) stage1_culled(S, k, G, C, o, d, r32, best, c);
    else stage1_brute(G, C, o, d, best, c);
    return seed_of(best);
}


// Stage 1 of one (ray, patch) pair by the whole wave (every lane calls it with
// the same arguments).  The pyramid is walked level by level, each level's
// boxes spread over the lanes (the caller's super mask ms_in, then the blocks
// of the hit supers 4 supers a pass, the tiles of the hit blocks 7 a pass, the
// cells of the hit tiles 7 a pass), and each hit cell's Moller-Trumbore tests
// run on a lane of its own.  The cells tested are the per-lane walk's, and the
// best hit is the minimum of (t, key) over them, so the result does not depend
// on the order; the work counts (returned; the super level's 9 box tests are
// the caller's) are the per-lane walk's too.  For a wave with one or two lanes
// left in the leaf phase, a level costs one memory round trip instead of one
// per hit box.
template <int CM>
__device__ __forceinline__ Stage1 stage1_coop(const DevScene& S, int64_t k, const V3& o, const V3& d, const Ray32& r, double widen,
                                              bool use_cull, uint32_t ms_in, uint32_t& n_cull, uint32_t& n_mt) {
    const int lane = __lane_id();
    const double* G = S.cxyz + (size_t)k * bbrt::kCoarseN * bbrt::kCoarseN * 3;
    const double* C = S.cell + (size_t)k * bbrt::kCells * bbrt::kCells * 4;
    Counters dummy = {};
    Stage1 best{ 0.0, 0, false };
    n_cull = 0;
    n_mt = 0;
    auto test_cell = [&](int i, int j) -> bool {   // false: the compact cell's exact box rejects the ray
        if constexpr (CM == 1) {
            const CellData cd = compact_cell_any(S, k, i, j);
            if (use_cull && !compact_cell_box(cd, widen, r)) return false;
            cell_tests_data(cd, i, j, o, d, best, dummy);
        } else {
            cell_tests(G, C, i, j, o, d, best, dummy);
        }
        return true;
    };
    if (!use_cull) {
        for (int base = 0; base < bbrt::kCells * bbrt::kCells; base += 64) {
            const int cell = base + lane;
            if (cell < bbrt::kCells * bbrt::kCells) test_cell(cell / bbrt::kCells, cell % bbrt::kCells);
        }
        n_mt = 6u * bbrt::kCells * bbrt::kCells;
    } else {
        const float* gb = S.box_block + (size_t)k * bbrt::kBlockFloats;
        const float* gt = S.box_tile + (size_t)k * bbrt::kTileFloats;
        const float* gc = S.box_cell + (size_t)k * bbrt::kCellFloats;
        const unsigned long long ms = ms_in;   // the super level, tested and counted by the caller
        const int ns = (int)__popcll(ms);
        for (int sb = 0; sb < ns; sb += 4) {
            // blocks of hit supers sb .. sb+3: lane = 16 * super + block slot
            const int sk = sb + (lane >> 4);
            const int sq = nth_bit(ms, min(sk, ns - 1));
            const int bq = lane & 15;
            const int blk = (bbrt::kBlocksPerSuper * (sq / 3) + (bq >> 2)) * bbrt::kBlockN +
                            bbrt::kBlocksPerSuper * (sq % 3) + (bq & 3);
            const bool h1 = sk < ns && box32(gb + sq * bbrt::kG16 + 6 * bq, r);   // empty slots never pass
            n_cull += 16u * (uint32_t)min(4, ns - sb);
            const unsigned long long m1 = __ballot(h1);
            const int nb = (int)__popcll(m1);
            for (int tb = 0; tb < nb; tb += 7) {
                // tiles of hit blocks tb .. tb+6: lane = 9 * block + tile slot
                const int bk = tb + lane / 9;
                const int b1 = __shfl(blk, nth_bit(m1, min(bk, nb - 1)));
                const int tq = lane % 9;
                const int bi = b1 / bbrt::kBlockN, bj = b1 % bbrt::kBlockN;
                const int tile = (3 * bi + tq / 3) * bbrt::kTileN + 3 * bj + tq % 3;
                const bool h2 = lane < 63 && bk < nb && box32(gt + b1 * bbrt::kG9 + 6 * tq, r);
                n_cull += 9u * (uint32_t)min(7, nb - tb);
                const unsigned long long m2 = __ballot(h2);
                const int nt = (int)__popcll(m2);
                for (int cb = 0; cb < nt; cb += 7) {
                    // cells of hit tiles cb .. cb+6: lane = 9 * tile + cell slot
                    const int tk = cb + lane / 9;
                    const int t1 = __shfl(tile, nth_bit(m2, min(tk, nt - 1)));
                    const int cq = lane % 9;
                    const int ti = t1 / bbrt::kTileN, tj = t1 % bbrt::kTileN;
                    const int cell = (3 * ti + cq / 3) * bbrt::kCells + 3 * tj + cq % 3;
                    bool h3 = false;
                    if (lane < 63 && tk < nt) {
                        if constexpr (CM == 1) {
                            // the quantized box of the cell, relative to its tile's box
                            const int tslot = 3 * (ti % 3) + tj % 3;
                            const float* tb6 = gt + ((ti / 3) * bbrt::kBlockN + tj / 3) * bbrt::kG9 + 6 * tslot;
                            const uint8_t* qc = S.qbox_cell + (size_t)k * bbrt::kQCellPatchBytes +
                                                (size_t)t1 * bbrt::kQCellBytes + 6 * cq;
                            const double tlo[3] = { tb6[0], tb6[1], tb6[2] };
                            const double step[3] = { ((double)tb6[3] - tlo[0]) / 255.0, ((double)tb6[4] - tlo[1]) / 255.0,
                                                     ((double)tb6[5] - tlo[2]) / 255.0 };
                            const float bx[6] = { round_down_f(tlo[0] + (double)qc[0] * step[0]),
                                                  round_down_f(tlo[1] + (double)qc[1] * step[1]),
                                                  round_down_f(tlo[2] + (double)qc[2] * step[2]),
                                                  round_up_f(tlo[0] + (double)qc[3] * step[0]),
                                                  round_up_f(tlo[1] + (double)qc[4] * step[1]),
                                                  round_up_f(tlo[2] + (double)qc[5] * step[2]) };
                            h3 = box32(bx, r);
                        } else {
                            h3 = box32(gc + t1 * bbrt::kG9 + 6 * cq, r);
                        }
                    }
                    n_cull += 9u * (uint32_t)min(7, nt - cb);
                    const unsigned long long m3 = __ballot(h3);
                    const int nc = (int)__popcll(m3);
                    // one hit cell per lane (nc <= 63)
                    const int c1 = __shfl(cell, nth_bit(m3, min(lane, nc - 1)));
                    bool ran = false;
                    if (lane < nc) ran = test_cell(c1 / bbrt::kCells, c1 % bbrt::kCells);
                    n_mt += 6u * (uint32_t)__popcll(__ballot(ran));
                }
            }
        }
    }
    // the wave's best (t, key)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double t2 = __shfl_xor(best.t, off);
        const int k2 = __shfl_xor(best.key, off);
        const bool h2 = __shfl_xor(best.have ? 1 : 0, off) != 0;
        if (h2 && (!best.have || t2 < best.t || (t2 == best.t && k2 < best.key))) { best.t = t2; best.key = k2; best.have = true; }
    }
    return best;
}

// Inclusive sum over the wave's lanes.
__device__ __forceinline__ int wave_incl_sum(int v) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    return v;
}

// The lane whose entries (inclusive sums `incl`) hold list entry e: the number of
// lanes whose entries all come before e.  Every lane calls it.
__device__ __forceinline__ int wave_find(int incl, int e) {
    int src = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
        if (__shfl(incl, src + step - 1) <= e) src += step;
    return src;
}

// Stage 1 of every leaf lane's (ray, patch) pair at once, with culling (every
// lane calls it; `own` marks the lanes with a pair, `sup` their hit super
// boxes, tested and counted by the caller).  The pyramid is walked
// breadth-first over the whole wave.  A pass gives each lane one entry of the
// current level, an (owner, parent box) pair, and the lane tests the parent's
// whole group of child boxes for the owner's ray in one memory round trip: the
// blocks of a hit super, the tiles of a hit block, the cells of a hit tile (CM
// 1, compact scenes: the tile's quantized cell boxes); the hit children are
// the next level's entries, and each hit cell's six Moller-Trumbore tests take
// one lane (CM 1: the cell evaluated from the fields, its exact box tested
// first).  Entries stay in owner order, so a segmented scan folds a pass's cell
// results per owner.  The boxes and cells tested are the per-lane walk's and
// each owner's best is the minimum of (t, key) over its cells, so the results
// and counts are the walk's; but a level costs one round trip per 64 entries
// instead of one per hit box of each lane.  `slots` is this wave's 64 owner
// records in LDS.
template <int CM>
__device__ __forceinline__ void stage1_bfs(const DevScene& S, bool own, int64_t k, const V3& o, const V3& d, const Ray32& r,
                                           double widen, uint32_t sup, Owner* slots, Stage1& best, uint32_t& n_cull,
                                           uint32_t& n_mt) {
    const int lane = __lane_id();
    if (own) {
        Owner& w = slots[lane];
        w.ox = o.x; w.oy = o.y; w.oz = o.z; w.dx = d.x; w.dy = d.y; w.dz = d.z;
        w.t = 0.0;
        w.widen = widen;
        w.r32[0] = r.ox; w.r32[1] = r.oy; w.r32[2] = r.oz; w.r32[3] = r.ix; w.r32[4] = r.iy; w.r32[5] = r.iz;
        w.k = k;
        w.use_cull = 1;
        w.n_cull = 16u * (uint32_t)__popc(sup);   // the blocks of each hit super
        w.n_mt = 0;
        w.key = 0; w.have = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    auto ray_of = [&](int ow) -> Ray32 {
        const Owner& w = slots[ow];
        return Ray32{ w.r32[0], w.r32[1], w.r32[2], w.r32[3], w.r32[4], w.r32[5] };
    };
    // level 1: (owner, hit super) -> the super's 16 block boxes
    const int c1 = own ? __popc(sup) : 0;
    const int i1 = wave_incl_sum(c1);
    const int n1 = __shfl(i1, 63);
    for (int p1 = 0; p1 < n1; p1 += 64) {
        const int e1 = p1 + lane;
        const int o1 = wave_find(i1, e1);
        const uint32_t s1 = (uint32_t)__shfl((int)sup, o1);
        const int b1 = __shfl(i1 - c1, o1);
        const bool v1 = e1 < n1;
        int sq = 0;
        uint32_t m1 = 0;
        if (v1) {
            sq = nth_bit(s1, e1 - b1);
            m1 = group_mask<16, bbrt::kG16>(S.box_block + (size_t)slots[o1].k * bbrt::kBlockFloats + sq * bbrt::kG16,
                                            ray_of(o1));
            if (m1) atomicAdd(&slots[o1].n_cull, 9u * (uint32_t)__popc(m1));   // the tiles of each hit block
        }
        // level 2: (owner, hit block) -> the block's 9 tile boxes
        const int c2 = __popc(m1);
        const int i2 = wave_incl_sum(c2);
        const int n2 = __shfl(i2, 63);
        for (int p2 = 0; p2 < n2; p2 += 64) {
            const int e2 = p2 + lane;
            const int src2 = wave_find(i2, e2);
            const uint32_t mm1 = (uint32_t)__shfl((int)m1, src2);
            const int b2 = __shfl(i2 - c2, src2);
            const int sq2 = __shfl(sq, src2);
            const int o2 = __shfl(o1, src2);
            const bool v2 = e2 < n2;
            int blk = 0;
            uint32_t m2 = 0;
            if (v2) {
                const int bq = nth_bit(mm1, e2 - b2);
                const int bi = bbrt::kBlocksPerSuper * (sq2 / 3) + (bq >> 2);
                const int bj = bbrt::kBlocksPerSuper * (sq2 % 3) + (bq & 3);
                blk = bi * bbrt::kBlockN + bj;
                m2 = group_mask<9, bbrt::kG9>(S.box_tile + (size_t)slots[o2].k * bbrt::kTileFloats + blk * bbrt::kG9,
                                              ray_of(o2));
                if (m2) atomicAdd(&slots[o2].n_cull, 9u * (uint32_t)__popc(m2));   // the cells of each hit tile
            }
            // level 3: (owner, hit tile) -> the tile's 9 cell boxes
            const int c3 = __popc(m2);
            const int i3 = wave_incl_sum(c3);
            const int n3 = __shfl(i3, 63);
            for (int p3 = 0; p3 < n3; p3 += 64) {
                const int e3 = p3 + lane;
                const int src3 = wave_find(i3, e3);
                const uint32_t mm2 = (uint32_t)__shfl((int)m2, src3);
                const int b3 = __shfl(i3 - c3, src3);
                const int blk3 = __shfl(blk, src3);
                const int o3 = __shfl(o2, src3);
                const bool v3 = e3 < n3;
                int tile = 0;
                uint32_t m3 = 0;
                if (v3) {
                    const int tq = nth_bit(mm2, e3 - b3);
                    const int ti = 3 * (blk3 / bbrt::kBlockN) + tq / 3, tj = 3 * (blk3 % bbrt::kBlockN) + tq % 3;
                    tile = ti * bbrt::kTileN + tj;
                    const int64_t k3 = slots[o3].k;
                    if constexpr (CM == 1) {
                        m3 = qcell_mask(S, k3, S.box_tile + (size_t)k3 * bbrt::kTileFloats + blk3 * bbrt::kG9 + 6 * tq,
                                        ti, tj, ray_of(o3));
                    } else {
                        m3 = group_mask<9, bbrt::kG9>(S.box_cell + (size_t)k3 * bbrt::kCellFloats + tile * bbrt::kG9,
                                                      ray_of(o3));
                        if (m3) atomicAdd(&slots[o3].n_mt, 6u * (uint32_t)__popc(m3));   // each hit cell's tests
                    }
                }
                // level 4: (owner, hit cell) -> its six Moller-Trumbore tests
                const int c4 = __popc(m3);
                const int i4 = wave_incl_sum(c4);
                const int n4 = __shfl(i4, 63);
                for (int p4 = 0; p4 < n4; p4 += 64) {
                    const int e4 = p4 + lane;
                    const int src4 = wave_find(i4, e4);
                    const uint32_t mm3 = (uint32_t)__shfl((int)m3, src4);
                    const int b4 = __shfl(i4 - c4, src4);
                    const int tile4 = __shfl(tile, src4);
                    const bool v4 = e4 < n4;
                    const int o4s = __shfl(o3, src4);   // every lane takes part in the shuffle
                    const int o4 = v4 ? o4s : -1;
                    Stage1 ib{ 0.0, 0, false };
                    if (v4) {
                        const int cq = nth_bit(mm3, e4 - b4);
                        const int ci = 3 * (tile4 / bbrt::kTileN) + cq / 3, cj = 3 * (tile4 % bbrt::kTileN) + cq % 3;
                        const int64_t kk = slots[o4].k;
                        Counters dummy = {};
                        if constexpr (CM == 1) {
                            const CellData cd = compact_cell_any(S, kk, ci, cj);
                            // the owner's ray only now: it is not held across the cell's field loads
                            __builtin_amdgcn_sched_barrier(0);
                            const Owner& w = slots[o4];
                            if (compact_cell_box(cd, w.widen, ray_of(o4))) {
                                const V3 oo = { w.ox, w.oy, w.oz }, dd = { w.dx, w.dy, w.dz };
                                cell_tests_data(cd, ci, cj, oo, dd, ib, dummy);
                                atomicAdd(&slots[o4].n_mt, 6u);
                            }
                        } else {
                            const Owner& w = slots[o4];
                            const V3 oo = { w.ox, w.oy, w.oz }, dd = { w.dx, w.dy, w.dz };
                            cell_tests(S.cxyz + (size_t)kk * bbrt::kCoarseN * bbrt::kCoarseN * 3,
                                       S.cell + (size_t)kk * bbrt::kCells * bbrt::kCells * 4, ci, cj, oo, dd, ib, dummy);
                        }
                    }
                    // per owner: the minimum (t, key) of its entries, at its last entry of the pass
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const double t2 = __shfl_up(ib.t, off);
                        const int k2 = __shfl_up(ib.key, off);
                        const bool h2 = __shfl_up(ib.have ? 1 : 0, off) != 0;
                        const int ow2 = __shfl_up(o4, off);
                        if (lane >= off && ow2 == o4 && h2 && (!ib.have || t2 < ib.t || (t2 == ib.t && k2 < ib.key))) {
                            ib.t = t2; ib.key = k2; ib.have = true;
                        }
                    }
                    const int next_ow = __shfl_down(o4, 1);
                    if (v4 && ib.have && (lane == 63 || next_ow != o4)) {
                        Owner& w = slots[o4];
                        if (!w.have || ib.t < w.t || (ib.t == w.t && ib.key < w.key)) {
                            w.t = ib.t; w.key = ib.key; w.have = 1;
                        }
                    }
                    // the next pass's owners read these records
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (own) {
        const Owner& w = slots[lane];
        best = Stage1{ w.t, w.key, w.have != 0 };
        n_cull = w.n_cull;
        n_mt = w.n_mt;
    }
    // the records are written again by the next call: every lane's reads of this call come first
    __builtin_amdgcn_wave_barrier();
}

// The triangle-centroid seed of a stage-1 hit (SurfacePatch.hpp:298-301).
__device__ __forceinline__ Seed seed_of(const Stage1& best) {
    Seed sd{ 0.0, 0.0, best.t, best.have };
    if (!best.have) return sd;
    const int cellk = best.key / 6;
    const int i = cellk / bbrt::kCells, j = cellk % bbrt::kCells;
    const bool second = (best.key & 1) != 0;
    const double ui = -1.0 + 2.0 * static_cast<double>(i) / (bbrt::kCoarseN - 1);
    const double uip1 = -1.0 + 2.0 * static_cast<double>(i + 1) / (bbrt::kCoarseN - 1);
    const double vj = -1.0 + 2.0 * static_cast<double>(j) / (bbrt::kCoarseN - 1);
    const double vjp1 = -1.0 + 2.0 * static_cast<double>(j + 1) / (bbrt::kCoarseN - 1);
    if (!second) { sd.u = (ui + uip1 + uip1) / 3.0; sd.v = (vj + vj + vjp1) / 3.0; }
    else { sd.u = (ui + ui + uip1) / 3.0; sd.v = (vj + vjp1 + vjp1) / 3.0; }
    return sd;
}

// newton_refine from (u0, v0, t0) on patch k, specialised on the node count.
template <bool F32 = false, bool PROBE = false>
__device__ __forceinline__ PHit newton_any(const DevScene& S, int64_t k, const V3& o, const V3& d, double u0, double v0,
                                           double t0, int mode, Counters& c) {
    const PatchMeta M = S.meta[k];
    if (M.imax == 3 && M.jmax == 3) return newton<3, F32, PROBE>(S, M, o, d, u0, v0, t0, mode, c);
    if (M.imax == 5 && M.jmax == 5) return newton<5, F32, PROBE>(S, M, o, d, u0, v0, t0, mode, c);
    return newton<0, F32, PROBE>(S, M, o, d, u0, v0, t0, mode, c);
}

// SurfacePatch::intersect (SurfacePatch.hpp:253-303): stage 1, then Newton from its seed.
__device__ PHit patch_intersect(const DevScene& S, int64_t k, const V3& o, const V3& d, const Ray32& r32, bool cull,
                                double widen, int mode, Counters& c) {
    const Seed sd = stage1_seed<2>(S, k, o, d, r32, cull, widen, c);
    if (!sd.have) return PHit{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    return newton_any(S, k, o, d, sd.u, sd.v, sd.t, mode, c);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void flush_counters(const Counters& c, unsigned long long* ctr) {
    const unsigned long long v[15] = { c.bvh_tests, c.candidates, c.cull_tests, c.mt_tests, c.newton_calls,
                                       c.newton_iters, c.newton_flops, c.polish_fail, c.bounces, c.seed_learned,
                                       c.seed_fallback, c.mlp_evals, c.mlp_lanes, c.newton_iters32,
                                       c.seed_miss };  // widened
    const int slot[15] = { bbrt::kCtrBvh, bbrt::kCtrCand, bbrt::kCtrCull, bbrt::kCtrMt, bbrt::kCtrNewtonCalls,
                           bbrt::kCtrNewtonIters, bbrt::kCtrNewtonFlops, bbrt::kCtrPolishFail, bbrt::kCtrBounces,
                           BBRT_CTR_SEED_LEARNED, BBRT_CTR_SEED_FALLBACK, BBRT_CTR_MLP_EVALS, BBRT_CTR_MLP_LANES,
                           BBRT_CTR_NEWTON_ITERS_FP32, BBRT_CTR_SEED_MISS };
#pragma unroll
    for (int q = 0; q < 15; ++q) {
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
__device__ __forceinline__ bool box_static6(const V3& o, const double inv[3], const double lo[3], const double hi[3],
                                            double& tmin_out) {
    double tmin = -INFINITY, tmax = INFINITY;
    const double oo[3] = { o.x, o.y, o.z };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t1 = (lo[a] - oo[a]) * inv[a];
        double t2 = (hi[a] - oo[a]) * inv[a];
        if (inv[a] < 0.0) { const double s = t1; t1 = t2; t2 = s; }
        tmin = smax(tmin, t1);
        tmax = smin(tmax, t2);
        if (tmax < tmin) return false;
    }
    tmin_out = tmin;
    return tmax >= smax(tmin, 0.0);
}
__device__ __forceinline__ bool box_static(const V3& o, const double inv[3], const BvhNode& nd, double& tmin_out) {
    return box_static6(o, inv, nd.lo, nd.hi, tmin_out);
}

// Traversal stack entry: idx < 0 marks a node whose box test already failed;
// meta >= 0 is an internal node (right child << 2 | split axis), meta < 0 a
// leaf (-1 - (start << 3 | count)); tmin is the node's entry distance.
// 4-wide traversal stack entry: node or leaf reference and its FP32 entry t.
struct Ent4 {
    int32_t ref;
    float tmin;
};

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
__device__ int traverse(const DevScene& S, int32_t root_idx, double widen, const V3& o, const V3& d, const Ray32& r32,
                        bool cull, int mode, bool root_live, PHit& best, int32_t& best_obj, int32_t& err_obj,
                        Counters& c) {
    const double inv[3] = { 1.0 / d.x, 1.0 / d.y, 1.0 / d.z };
    const bool neg[3] = { inv[0] < 0.0, inv[1] < 0.0, inv[2] < 0.0 };
    double best_t = INFINITY;
    best_obj = -1;
    err_obj = -1;
    Ent stack[kStackCap];
    int sp = 0;
    Ent cur;
    {
        const BvhNode root = S.bvh[root_idx];
        double tm = -INFINITY;
        const bool ok = root_live ? true : box_static(o, inv, root, tm);
        cur = make_ent(root_idx, root, ok, tm);
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
                const PHit h = patch_intersect(S, obj, o, d, r32, cull, widen, mode, c);
                if (h.status < 0) { err_obj = obj; return -h.status; }
                if (h.status == 1 && h.t < best_t) { best_t = h.t; best = h; best_obj = obj; }
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return 0;
}

// The end of one bounce, RayTracingDriver.hpp:190-237: error, escape, back
// face, or record + reflect + nudge.  Returns true while the ray continues;
// o and d are advanced in place.  Patch indices are recorded relative to the
// ray's sub-scene; hit_count is indexed by the global patch index.
__device__ __forceinline__ bool finish_bounce(const DevScene& S, const Outs& O, int32_t r, int scn, V3& o, V3& d, int bounce,
                                              int max_bounces, int record_k, uint32_t flags, int err, int32_t err_obj,
                                              const PHit& best, int32_t best_obj, Counters& c) {
    const bbrt::SceneInfo& SI = S.scenes[scn];
    if (err) {
        if (O.status) O.status[r] = bbrt::kError;
        if (O.err_kind) O.err_kind[r] = (int8_t)err;
        if (O.err_patch) O.err_patch[r] = err_obj >= 0 ? err_obj - SI.patch_begin : err_obj;
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
        if (O.err_patch) O.err_patch[r] = best_obj - SI.patch_begin;
        if (O.bounces) O.bounces[r] = bounce;
        write_final(O, r, o, d);
        return false;
    }
    if (bounce < record_k) {
        const size_t at = (size_t)r * record_k + bounce;
        if (O.hit_patch) O.hit_patch[at] = best_obj - SI.patch_begin;
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
    const double eps = smax(kSelfEpsFactor * SI.diag, kSagittaClearance * S.meta[best_obj].max_sag);
    o = { best.p.x + d.x * eps, best.p.y + d.y * eps, best.p.z + d.z * eps };
    if (bounce + 1 >= max_bounces) {
        if (O.status) O.status[r] = bbrt::kTrapped;
        if (O.bounces) O.bounces[r] = bounce + 1;
        write_final(O, r, o, d);
        return false;
    }
    return true;
}

__device__ __forceinline__ bool cull_ok(const DevScene& S, int scn, const V3& o, const V3& d, uint32_t flags) {
    const bool finite = isfinite(o.x) && isfinite(o.y) && isfinite(o.z) && isfinite(d.x) && isfinite(d.y) &&
                        isfinite(d.z);
    const double lim = S.scenes[scn].cull_limit;
    return !(flags & BBRT_FLAG_NO_CULL) && finite && fabs(o.x) <= lim && fabs(o.y) <= lim && fabs(o.z) <= lim;
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
    const int scn = ray_scene_of(S, r);
    const bool cull = cull_ok(S, scn, o, d, flags);
    const Ray32 r32 = { (float)o.x, (float)o.y, (float)o.z, 1.0f / (float)d.x, 1.0f / (float)d.y, 1.0f / (float)d.z };
    PHit best{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    int32_t best_obj, err_obj;
    const int err = traverse(S, S.scenes[scn].root, S.scenes[scn].cull_widen, o, d, r32, cull, mode, root_live, best,
                             best_obj, err_obj, c);
    const bool alive = finish_bounce(S, O, r, scn, o, d, bounce, max_bounces, record_k, flags, err, err_obj, best,
                                     best_obj, c);
    if constexpr (WORK) add_ray_work(O, r, c, c0);
    return alive;
}

}  // namespace
