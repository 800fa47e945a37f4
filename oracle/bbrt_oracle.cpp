// bbrt_oracle.cpp -- CPU FP64 oracle for the billiard bounce loop.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithm of
// manuelarturosantana/BillardBallRayTracing3D on the CPU so that the HIP path
// can be checked against it.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product library never does.
//
// Every function cites the reference file:line it follows.  Build flags:
// -O2 -ffp-contract=off, no -march: x86-64 baseline SSE2 with no FMA, like the
// reference's Release build (CMakeLists.txt:7-11).
//
// Pinning status of each part (what fixes its results besides this text):
//   moller_trumbore, ray_box, BVH build/traversal ... compiled reference
//       headers (oracle/_ref) and the fixtures written from them;
//   barycentric weights, first/second-form Lagrange interpolation ... closed
//       forms (polynomial reproduction, one-hot at nodes) in tests;
//   coarse grid ... polynomial reproduction; the MKL dgemm summation order of
//       PatchInterpolation.hpp:148-149 is NOT reproduced (parity unpinned at
//       that boundary: MKL is not in this image);
//   cell padding ... paraboloid closed form (sagitta = |a+b| (1/99)^2);
//   Newton ... paraboloid closed form, forced singular / no-convergence /
//       domain events (tests/test_oracle_math.py);
//   stage-1 winner and triangle_uv_centroid seed ... the coarse triangle
//       under a flat patch's hit point and the mean of its corner (u, v);
//       Newton from exactly that seed (tests/test_oracle_bounce.py);
//   bounce loop ... parallel flat plates (exact reflected directions, hit
//       points, bounce count, nudge 1e-9 * diagonal), reflection invariants
//       and the nudge 5 * max sagitta of the hit patch on paraboloid patches,
//       zero direction, immediate escape, head-on sphere, the periodic x-axis
//       orbit between two spheres (bit-exact, Trapped) and its divergence
//       rate 2 + sqrt(3) (tests/test_oracle_bounce.py).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <optional>
#include <vector>
#include <omp.h>

namespace {

constexpr double kEqualTol = 1.0e-12;          // PatchInterpolation.hpp:21
constexpr int kCoarse = 100;                    // SurfacePatch.hpp:117
constexpr int kNewtonMaxIters = 20;             // SurfacePatch.hpp:121
constexpr double kNewtonTol = 1e-3;             // SurfacePatch.hpp:122
constexpr double kDomainTol = 1e-6;             // SurfacePatch.hpp:639
constexpr double kDetTol = 1e-14;               // SurfacePatch.hpp:567
constexpr double kSelfEpsFactor = 1e-9;         // RayTracingDriver.hpp:328
constexpr double kSagittaClearance = 5.0;       // RayTracingDriver.hpp:337
constexpr double kBackfaceTol = 1e-9;           // RayTracingDriver.hpp:344
constexpr int kLeafThreshold = 4;               // BVH.hpp:67
// Polish of the `converged` Newton mode (no reference counterpart).
constexpr int kPolishMax = 8;
constexpr double kPolishRelTol = 1e-13;

enum ErrKind : int8_t { kErrNone = 0, kErrSingular = 1, kErrNoConv = 2, kErrDomain = 3,
                        kErrBackface = 4, kErrZeroDir = 5, kErrStack = 6 };
enum Status : int8_t { kEscaped = 0, kTrapped = 1, kError = 2 };

struct Vec3 { double x, y, z; };                  // ScatObject.hpp:10-12
struct AABB {                                     // ScatObject.hpp:19-26
    Vec3 min{ std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
              std::numeric_limits<double>::infinity() };
    Vec3 max{ -std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity(),
              -std::numeric_limits<double>::infinity() };
};

// MollerTrumbore.hpp:13-27
inline Vec3 vsub(const Vec3& a, const Vec3& b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline Vec3 vcross(const Vec3& a, const Vec3& b) {
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
inline double vdot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 vnormalize(const Vec3& v) {
    double len = std::sqrt(vdot(v, v));
    return { v.x / len, v.y / len, v.z / len };
}

// MollerTrumbore.hpp:32-64.  Returns true on a hit; t, point and normal out.
inline bool moller_trumbore(const Vec3& o, const Vec3& d, const Vec3& v0, const Vec3& v1, const Vec3& v2,
                            double eps, double& t_out, Vec3* point, Vec3* normal) {
    Vec3 edge1 = vsub(v1, v0);
    Vec3 edge2 = vsub(v2, v0);
    Vec3 pvec = vcross(d, edge2);
    double det = vdot(edge1, pvec);
    if (std::fabs(det) < eps) return false;
    double inv_det = 1.0 / det;
    Vec3 tvec = vsub(o, v0);
    double u = vdot(tvec, pvec) * inv_det;
    if (u < 0.0 || u > 1.0) return false;
    Vec3 qvec = vcross(tvec, edge1);
    double v = vdot(d, qvec) * inv_det;
    if (v < 0.0 || u + v > 1.0) return false;
    double t = vdot(edge2, qvec) * inv_det;
    if (t < eps) return false;
    t_out = t;
    if (point) *point = { o.x + t * d.x, o.y + t * d.y, o.z + t * d.z };
    if (normal) *normal = vnormalize(vcross(edge1, edge2));
    return true;
}

// BVH.hpp:14-37.  std::max(a,b) is (a < b) ? b : a and std::min(a,b) is
// (b < a) ? b : a, so a NaN second argument is ignored.
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline double smin(double a, double b) { return (b < a) ? b : a; }

inline bool ray_box(const Vec3& origin, const double inv_dir[3], const AABB& box, double t_hit_so_far,
                    double& tmin_out) {
    double tmin = -std::numeric_limits<double>::infinity();
    double tmax = std::numeric_limits<double>::infinity();
    const double o[3] = { origin.x, origin.y, origin.z };
    const double bmin[3] = { box.min.x, box.min.y, box.min.z };
    const double bmax[3] = { box.max.x, box.max.y, box.max.z };
    for (int axis = 0; axis < 3; ++axis) {
        double t1 = (bmin[axis] - o[axis]) * inv_dir[axis];
        double t2 = (bmax[axis] - o[axis]) * inv_dir[axis];
        if (inv_dir[axis] < 0.0) std::swap(t1, t2);
        tmin = smax(tmin, t1);
        tmax = smin(tmax, t2);
        if (tmax < tmin) return false;
    }
    if (tmin > t_hit_so_far) return false;
    tmin_out = tmin;
    return tmax >= smax(tmin, 0.0);
}

// BVH.hpp:39-48
inline AABB union_box(const AABB& a, const AABB& b) {
    AABB r;
    r.min = { smin(a.min.x, b.min.x), smin(a.min.y, b.min.y), smin(a.min.z, b.min.z) };
    r.max = { smax(a.max.x, b.max.x), smax(a.max.y, b.max.y), smax(a.max.z, b.max.z) };
    return r;
}

// PatchInterpolation.hpp:25-52
std::vector<double> barycentric_weights(const double* nodes, int N) {
    std::vector<double> w(N, 1.0);
    for (int j = 0; j < N; j++) {
        double p = 1.0;
        for (int i = 0; i < N; i++)
            if (i != j) p *= (nodes[j] - nodes[i]);
        w[j] = 1.0 / p;
    }
    return w;
}

// PatchInterpolation.hpp:54-159: first-form barycentric interpolation on a
// P x Q tensor grid.  The two cblas_dgemm calls (:148-149) are restated as
// plain loops summing k (then m) upwards from 0.0; MKL's own order is not
// reproduced.
void interp2d(const double* xN, int N, const double* yN, int M, const double* xW, const double* yW,
              const double* F, const double* x, int P, const double* y, int Q, double* sol) {
    std::vector<double> lx(P, 1.0), ly(Q, 1.0), vx((size_t)P * N, 0.0), vy((size_t)M * Q, 0.0);
    for (int i = 0; i < P; ++i) {
        double prod = 1.0; int eq = -1;
        for (int j = 0; j < N; ++j) {
            double diff = x[i] - xN[j];
            if (std::abs(diff) < kEqualTol) { eq = j; break; }
            vx[(size_t)i * N + j] = xW[j] / diff;
            prod *= diff;
        }
        if (eq != -1) {
            std::fill(vx.begin() + (size_t)i * N, vx.begin() + (size_t)(i + 1) * N, 0.0);
            vx[(size_t)i * N + eq] = 1.0; lx[i] = 1.0;
        } else lx[i] = prod;
    }
    for (int i = 0; i < Q; ++i) {
        double prod = 1.0; int eq = -1;
        for (int j = 0; j < M; ++j) {
            double diff = y[i] - yN[j];
            if (std::abs(diff) < kEqualTol) { eq = j; break; }
            vy[(size_t)i * M + j] = yW[j] / diff;
            prod *= diff;
        }
        if (eq != -1) {
            std::fill(vy.begin() + (size_t)i * M, vy.begin() + (size_t)(i + 1) * M, 0.0);
            vy[(size_t)i * M + eq] = 1.0; ly[i] = 1.0;
        } else ly[i] = prod;
    }
    std::vector<double> mp((size_t)P * M);
    for (int i = 0; i < P; ++i)
        for (int m = 0; m < M; ++m) {
            double acc = 0.0;
            for (int k = 0; k < N; ++k) acc += vx[(size_t)i * N + k] * F[(size_t)k * M + m];
            mp[(size_t)i * M + m] = acc;
        }
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < Q; ++j) {
            double acc = 0.0;
            for (int m = 0; m < M; ++m) acc += mp[(size_t)i * M + m] * vy[(size_t)j * M + m];
            sol[(size_t)i * Q + j] = acc;
        }
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < Q; ++j) sol[(size_t)i * Q + j] *= lx[i] * ly[j];
}

// SurfacePatch.hpp:518-538 (second-form basis, one-hot on a node)
void basis1d(const double* nodes, const double* weights, int N, double x, double* basis) {
    for (int i = 0; i < N; ++i) basis[i] = 0.0;
    for (int i = 0; i < N; ++i)
        if (std::fabs(x - nodes[i]) < kEqualTol) { basis[i] = 1.0; return; }
    double denom = 0.0;
    for (int i = 0; i < N; ++i) {
        double li = weights[i] / (x - nodes[i]);
        basis[i] = li;
        denom += li;
    }
    for (int i = 0; i < N; ++i) basis[i] /= denom;
}

// SurfacePatch.hpp:542-555
double eval_field(const double* lu, const double* lv, const double* f, int imax, int jmax) {
    double sum = 0.0;
    for (int i = 0; i < imax; ++i) {
        const double li = lu[i];
        if (li == 0.0) continue;
        const double* row = f + (size_t)i * jmax;
        double row_sum = 0.0;
        for (int j = 0; j < jmax; ++j) row_sum += lv[j] * row[j];
        sum += li * row_sum;
    }
    return sum;
}

// SurfacePatch.hpp:559-579
bool solve3x3(const double J[3][3], const Vec3& b, double x[3]) {
    const double a11 = J[0][0], a12 = J[0][1], a13 = J[0][2];
    const double a21 = J[1][0], a22 = J[1][1], a23 = J[1][2];
    const double a31 = J[2][0], a32 = J[2][1], a33 = J[2][2];
    const double det = a11 * (a22 * a33 - a23 * a32) - a12 * (a21 * a33 - a23 * a31) + a13 * (a21 * a32 - a22 * a31);
    if (std::fabs(det) < kDetTol) return false;
    const double bx = b.x, by = b.y, bz = b.z;
    const double det0 = bx * (a22 * a33 - a23 * a32) - a12 * (by * a33 - a23 * bz) + a13 * (by * a32 - a22 * bz);
    const double det1 = a11 * (by * a33 - a23 * bz) - bx * (a21 * a33 - a23 * a31) + a13 * (a21 * bz - by * a31);
    const double det2 = a11 * (a22 * bz - by * a32) - a12 * (a21 * bz - by * a31) + bx * (a21 * a32 - a22 * a31);
    x[0] = det0 / det; x[1] = det1 / det; x[2] = det2 / det;
    return true;
}

struct Patch {
    int imax = 0, jmax = 0;
    std::vector<double> un, vn, uw, vw;
    std::vector<double> f[12];     // x y z dxdu dydu dzdu dxdv dydv dzdv nuX nuY nuZ
    std::vector<double> cu, cv, cx, cy, cz, sag;
    std::vector<Vec3> cnorm;
    AABB box; Vec3 centroid{ 0, 0, 0 };
    double max_sag = 0.0;

    Vec3 cpoint(int i, int j) const { size_t k = (size_t)i * kCoarse + j; return { cx[k], cy[k], cz[k] }; }

    // SurfacePatch.hpp:380-407
    void centroid_and_box() {
        const size_t n = (size_t)imax * jmax;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        AABB b;
        for (size_t i = 0; i < n; ++i) {
            const double px = f[0][i], py = f[1][i], pz = f[2][i];
            sx += px; sy += py; sz += pz;
            b.min.x = smin(b.min.x, px); b.min.y = smin(b.min.y, py); b.min.z = smin(b.min.z, pz);
            b.max.x = smax(b.max.x, px); b.max.y = smax(b.max.y, py); b.max.z = smax(b.max.z, pz);
        }
        const double inv_n = 1.0 / static_cast<double>(n);
        centroid = { sx * inv_n, sy * inv_n, sz * inv_n };
        box = b;
    }
    // SurfacePatch.hpp:416-436
    void coarse_grid() {
        const int n = kCoarse;
        cu.resize(n); cv.resize(n);
        for (int i = 0; i < n; ++i) {
            double t = (n == 1) ? 0.0 : -1.0 + 2.0 * static_cast<double>(i) / (n - 1);
            cu[i] = t; cv[i] = t;
        }
        cx.resize((size_t)n * n); cy.resize((size_t)n * n); cz.resize((size_t)n * n);
        interp2d(un.data(), imax, vn.data(), jmax, uw.data(), vw.data(), f[0].data(), cu.data(), n, cv.data(), n, cx.data());
        interp2d(un.data(), imax, vn.data(), jmax, uw.data(), vw.data(), f[1].data(), cu.data(), n, cv.data(), n, cy.data());
        interp2d(un.data(), imax, vn.data(), jmax, uw.data(), vw.data(), f[2].data(), cu.data(), n, cv.data(), n, cz.data());
    }
    // SurfacePatch.hpp:446-492
    void cell_padding() {
        const int n = kCoarse;
        const int cells = (n - 1) * (n - 1);
        sag.assign(cells, 0.0);
        cnorm.assign(cells, Vec3{ 0, 0, 0 });
        std::vector<double> lu(imax), lv(jmax);
        double ms = 0.0;
        for (int i = 0; i + 1 < n; ++i)
            for (int j = 0; j + 1 < n; ++j) {
                const double um = 0.5 * (cu[i] + cu[i + 1]);
                const double vm = 0.5 * (cv[j] + cv[j + 1]);
                basis1d(un.data(), uw.data(), imax, um, lu.data());
                basis1d(vn.data(), vw.data(), jmax, vm, lv.data());
                Vec3 tm = { eval_field(lu.data(), lv.data(), f[0].data(), imax, jmax),
                            eval_field(lu.data(), lv.data(), f[1].data(), imax, jmax),
                            eval_field(lu.data(), lv.data(), f[2].data(), imax, jmax) };
                Vec3 nm = vnormalize({ eval_field(lu.data(), lv.data(), f[9].data(), imax, jmax),
                                       eval_field(lu.data(), lv.data(), f[10].data(), imax, jmax),
                                       eval_field(lu.data(), lv.data(), f[11].data(), imax, jmax) });
                const Vec3 p00 = cpoint(i, j), p10 = cpoint(i + 1, j), p01 = cpoint(i, j + 1), p11 = cpoint(i + 1, j + 1);
                const Vec3 bm = { 0.25 * (p00.x + p10.x + p01.x + p11.x), 0.25 * (p00.y + p10.y + p01.y + p11.y),
                                  0.25 * (p00.z + p10.z + p01.z + p11.z) };
                const Vec3 df = vsub(tm, bm);
                const double s = std::sqrt(vdot(df, df));
                const int idx = i * (n - 1) + j;
                sag[idx] = s; cnorm[idx] = nm;
                ms = smax(ms, s);
            }
        box.min.x -= ms; box.min.y -= ms; box.min.z -= ms;
        box.max.x += ms; box.max.y += ms; box.max.z += ms;
        max_sag = ms;
    }
};

struct Counters {
    int64_t newton_calls = 0, newton_iters = 0, candidates = 0, mt_tests = 0, box_tests = 0, polish_fail = 0;
};

// Result of one patch intersection: status 1 = hit, 0 = miss, <0 = -ErrKind.
struct PatchHit { int status; double t; Vec3 point; Vec3 normal; int iters; double u = 0.0, v = 0.0; };

// SurfacePatch.hpp:593-660 (+ the `converged` polish, which has no counterpart)
PatchHit newton_refine(const Patch& P, const Vec3& o, const Vec3& d, double u0, double v0, double t0, int mode,
                       Counters& C) {
    PatchHit r{ 0, 0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    double u = u0, v = v0, t = t0;
    double lu[32], lv[32];
    double step_norm = std::numeric_limits<double>::infinity();
    const int im = P.imax, jm = P.jmax;
    C.newton_calls++;
    auto iterate = [&](double delta[3]) -> bool {
        basis1d(P.un.data(), P.uw.data(), im, u, lu);
        basis1d(P.vn.data(), P.vw.data(), jm, v, lv);
        Vec3 X = { eval_field(lu, lv, P.f[0].data(), im, jm), eval_field(lu, lv, P.f[1].data(), im, jm),
                   eval_field(lu, lv, P.f[2].data(), im, jm) };
        Vec3 Xu = { eval_field(lu, lv, P.f[3].data(), im, jm), eval_field(lu, lv, P.f[4].data(), im, jm),
                    eval_field(lu, lv, P.f[5].data(), im, jm) };
        Vec3 Xv = { eval_field(lu, lv, P.f[6].data(), im, jm), eval_field(lu, lv, P.f[7].data(), im, jm),
                    eval_field(lu, lv, P.f[8].data(), im, jm) };
        Vec3 R = { o.x + t * d.x, o.y + t * d.y, o.z + t * d.z };
        Vec3 F = vsub(X, R);
        const double J[3][3] = { { Xu.x, Xv.x, -d.x }, { Xu.y, Xv.y, -d.y }, { Xu.z, Xv.z, -d.z } };
        return solve3x3(J, F, delta);
    };
    int iters = 0;
    for (int iter = 0; iter < kNewtonMaxIters; ++iter) {
        double delta[3];
        if (!iterate(delta)) { C.newton_iters += iters; r.status = -kErrSingular; r.iters = iters; return r; }
        u -= delta[0]; v -= delta[1]; t -= delta[2];
        ++iters;
        step_norm = std::sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
        if (step_norm < kNewtonTol) break;
    }
    if (step_norm >= kNewtonTol) { C.newton_iters += iters; r.status = -kErrNoConv; r.iters = iters; return r; }
    if (u < -1.0 - kDomainTol || u > 1.0 + kDomainTol || v < -1.0 - kDomainTol || v > 1.0 + kDomainTol) {
        C.newton_iters += iters; r.status = -kErrDomain; r.iters = iters; return r;
    }
    if (mode == 1) {
        bool ok = false;
        for (int k = 0; k < kPolishMax; ++k) {
            double delta[3];
            if (!iterate(delta)) break;
            u -= delta[0]; v -= delta[1]; t -= delta[2];
            ++iters;
            const double sn = std::sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
            if (sn < kPolishRelTol * smax(1.0, std::fabs(t))) { ok = true; break; }
        }
        if (!ok) C.polish_fail++;
    }
    C.newton_iters += iters;
    basis1d(P.un.data(), P.uw.data(), im, u, lu);
    basis1d(P.vn.data(), P.vw.data(), jm, v, lv);
    r.status = 1; r.t = t; r.iters = iters; r.u = u; r.v = v;
    r.point = { eval_field(lu, lv, P.f[0].data(), im, jm), eval_field(lu, lv, P.f[1].data(), im, jm),
                eval_field(lu, lv, P.f[2].data(), im, jm) };
    r.normal = vnormalize({ eval_field(lu, lv, P.f[9].data(), im, jm), eval_field(lu, lv, P.f[10].data(), im, jm),
                            eval_field(lu, lv, P.f[11].data(), im, jm) });
    return r;
}

// Stage 1 of SurfacePatch::intersect (SurfacePatch.hpp:253-298): best (t, i, j, second).
bool stage1(const Patch& P, const Vec3& o, const Vec3& d, double& bt, int& bi, int& bj, bool& bsec, Counters& C) {
    const int n = kCoarse;
    bool have = false;
    C.candidates++;
    for (int i = 0; i + 1 < n; ++i)
        for (int j = 0; j + 1 < n; ++j) {
            Vec3 p00 = P.cpoint(i, j), p10 = P.cpoint(i + 1, j), p01 = P.cpoint(i, j + 1), p11 = P.cpoint(i + 1, j + 1);
            const int idx = i * (n - 1) + j;
            const double s = P.sag[idx];
            const Vec3& cn = P.cnorm[idx];
            for (double sign : { 0.0, 1.0, -1.0 }) {
                const double dd = sign * s;
                Vec3 q00 = { p00.x + cn.x * dd, p00.y + cn.y * dd, p00.z + cn.z * dd };
                Vec3 q10 = { p10.x + cn.x * dd, p10.y + cn.y * dd, p10.z + cn.z * dd };
                Vec3 q01 = { p01.x + cn.x * dd, p01.y + cn.y * dd, p01.z + cn.z * dd };
                Vec3 q11 = { p11.x + cn.x * dd, p11.y + cn.y * dd, p11.z + cn.z * dd };
                double t;
                if (moller_trumbore(o, d, q00, q10, q11, 1e-9, t, nullptr, nullptr) && (!have || t < bt)) {
                    have = true; bt = t; bi = i; bj = j; bsec = false;
                }
                if (moller_trumbore(o, d, q00, q11, q01, 1e-9, t, nullptr, nullptr) && (!have || t < bt)) {
                    have = true; bt = t; bi = i; bj = j; bsec = true;
                }
            }
        }
    C.mt_tests += 6LL * (n - 1) * (n - 1);
    return have;
}

// triangle_uv_centroid (SurfacePatch.hpp:499-510): the (u, v) seed of the stage-1
// winner, triangle (p00, p10, p11) or, when `second`, (p00, p11, p01).
void triangle_uv_centroid(const Patch& P, int i, int j, bool second, double& u0, double& v0) {
    const double ui = P.cu[i], uip1 = P.cu[i + 1], vj = P.cv[j], vjp1 = P.cv[j + 1];
    if (!second) { u0 = (ui + uip1 + uip1) / 3.0; v0 = (vj + vj + vjp1) / 3.0; }
    else { u0 = (ui + ui + uip1) / 3.0; v0 = (vj + vjp1 + vjp1) / 3.0; }
}

// SurfacePatch.hpp:253-303 (+ triangle_uv_centroid, :499-510)
PatchHit patch_intersect(const Patch& P, const Vec3& o, const Vec3& d, int mode, Counters& C) {
    double bt = 0; int bi = -1, bj = -1; bool bsec = false;
    if (!stage1(P, o, d, bt, bi, bj, bsec, C)) return PatchHit{ 0, 0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    double u0, v0;
    triangle_uv_centroid(P, bi, bj, bsec, u0, v0);
    return newton_refine(P, o, d, u0, v0, bt, mode, C);
}

struct Node { AABB box; int32_t right = -1, start = 0, count = 0; uint8_t axis = 0; };

// BVH.hpp:136-189, objects being indices into `boxes`/`cents`.
struct BVHBuild {
    const std::vector<AABB>& boxes; const std::vector<Vec3>& cents; int leaf_threshold;
    std::vector<Node> nodes; std::vector<int32_t> leaf;
    int32_t build(std::vector<int32_t> objs) {
        int32_t me = (int32_t)nodes.size();
        nodes.emplace_back();
        AABB box = boxes[objs.front()];
        for (size_t i = 1; i < objs.size(); ++i) box = union_box(box, boxes[objs[i]]);
        if ((int)objs.size() <= leaf_threshold) {
            nodes[me].box = box; nodes[me].right = -1;
            nodes[me].start = (int32_t)leaf.size(); nodes[me].count = (int32_t)objs.size();
            for (int32_t o : objs) leaf.push_back(o);
            return me;
        }
        Vec3 cmin = cents[objs.front()], cmax = cmin;
        for (int32_t o : objs) {
            const Vec3& c = cents[o];
            cmin.x = smin(cmin.x, c.x); cmax.x = smax(cmax.x, c.x);
            cmin.y = smin(cmin.y, c.y); cmax.y = smax(cmax.y, c.y);
            cmin.z = smin(cmin.z, c.z); cmax.z = smax(cmax.z, c.z);
        }
        double spread[3] = { cmax.x - cmin.x, cmax.y - cmin.y, cmax.z - cmin.z };
        uint8_t axis = (uint8_t)std::distance(spread, std::max_element(spread, spread + 3));
        auto mid = objs.begin() + objs.size() / 2;
        const std::vector<Vec3>& cs = cents;
        std::nth_element(objs.begin(), mid, objs.end(), [axis, &cs](int32_t a, int32_t b) {
            const Vec3& ca = cs[a]; const Vec3& cb = cs[b];
            double va = (axis == 0) ? ca.x : (axis == 1) ? ca.y : ca.z;
            double vb = (axis == 0) ? cb.x : (axis == 1) ? cb.y : cb.z;
            return va < vb;
        });
        std::vector<int32_t> lo(objs.begin(), mid), hi(mid, objs.end());
        objs.clear(); objs.shrink_to_fit();
        build(std::move(lo));
        int32_t r = build(std::move(hi));
        nodes[me].box = box; nodes[me].right = r; nodes[me].count = 0; nodes[me].axis = axis;
        return me;
    }
};

struct Scene {
    std::vector<Patch> patches;
    std::vector<Node> nodes; std::vector<int32_t> leaf;
    double diag = 0.0, max_patch_sag = 0.0;
};

// BVH.hpp:75-128 generalised over the per-object intersection.  `isect`
// returns a PatchHit; an error status aborts the traversal like the throw at
// BVH.hpp:103 propagating out of BVH::intersect.
template <class F>
PatchHit bvh_intersect(const std::vector<Node>& nodes, const std::vector<int32_t>& leaf, const Vec3& o, const Vec3& d,
                       F&& isect, int32_t& best_obj, int32_t& err_obj, Counters& C) {
    double inv[3] = { 1.0 / d.x, 1.0 / d.y, 1.0 / d.z };
    bool neg[3] = { inv[0] < 0.0, inv[1] < 0.0, inv[2] < 0.0 };
    PatchHit best{ 0, 0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    double best_t = std::numeric_limits<double>::infinity();
    int32_t stack[64]; int sp = 0;
    stack[sp++] = 0;
    best_obj = -1; err_obj = -1;
    while (sp > 0) {
        int32_t idx = stack[--sp];
        const Node& nd = nodes[idx];
        double tmin;
        C.box_tests++;
        if (!ray_box(o, inv, nd.box, best_t, tmin)) continue;
        if (nd.right < 0) {
            for (int32_t i = 0; i < nd.count; ++i) {
                const int32_t obj = leaf[nd.start + i];
                PatchHit h = isect(obj);
                if (h.status < 0) { err_obj = obj; return h; }
                if (h.status == 1 && h.t < best_t) { best_t = h.t; best = h; best_obj = obj; }
            }
            continue;
        }
        int32_t left = idx + 1, right = nd.right;
        if (sp + 2 > 64) { PatchHit e{ -kErrStack, 0, { 0, 0, 0 }, { 0, 0, 0 }, 0 }; return e; }
        if (neg[nd.axis]) { stack[sp++] = left; stack[sp++] = right; }
        else { stack[sp++] = right; stack[sp++] = left; }
    }
    return best;
}

struct RayOut {
    int8_t* status; int8_t* err_kind; int32_t* bounces; int32_t* err_patch;
    int32_t* hit_patch; double* hit_xyz; uint8_t* iters; double* final_od; int32_t record_k;
};

// RayTracingDriver.hpp:172-242 (trace_ray) with normalize_dir (:398-401),
// reflect (:405-408) and self_intersect_epsilon (:260-265).
void trace_ray(const Scene& S, int64_t ray, const Vec3& origin, const Vec3& direction, int max_bounces, int mode,
               const RayOut& out, Counters& C) {
    const int K = out.record_k;
    for (int k = 0; k < K; ++k) {
        out.hit_patch[ray * K + k] = -1;
        out.iters[ray * K + k] = 0;
        for (int c = 0; c < 3; ++c) out.hit_xyz[(ray * K + k) * 3 + c] = 0.0;
    }
    out.bounces[ray] = 0; out.err_kind[ray] = kErrNone; out.err_patch[ray] = -1;
    double* fod = out.final_od + ray * 6;
    for (int c = 0; c < 6; ++c) fod[c] = 0.0;
    const double len2 = direction.x * direction.x + direction.y * direction.y + direction.z * direction.z;
    if (len2 < 1e-300) {
        out.status[ray] = kError; out.err_kind[ray] = kErrZeroDir;
        fod[0] = origin.x; fod[1] = origin.y; fod[2] = origin.z;
        fod[3] = direction.x; fod[4] = direction.y; fod[5] = direction.z;
        return;
    }
    Vec3 o = origin;
    const double len = std::sqrt(direction.x * direction.x + direction.y * direction.y + direction.z * direction.z);
    Vec3 d = { direction.x / len, direction.y / len, direction.z / len };
    int bounces = 0;
    auto save = [&]() { fod[0] = o.x; fod[1] = o.y; fod[2] = o.z; fod[3] = d.x; fod[4] = d.y; fod[5] = d.z; };
    for (int b = 0; b < max_bounces; ++b) {
        int32_t best_obj, err_obj;
        PatchHit h = bvh_intersect(S.nodes, S.leaf, o, d,
                                   [&](int32_t k) { return patch_intersect(S.patches[k], o, d, mode, C); },
                                   best_obj, err_obj, C);
        if (h.status < 0) {
            out.status[ray] = kError; out.err_kind[ray] = (int8_t)(-h.status); out.err_patch[ray] = err_obj;
            out.bounces[ray] = bounces; save(); return;
        }
        if (best_obj < 0) { out.status[ray] = kEscaped; out.bounces[ray] = bounces; save(); return; }
        if (vdot(d, h.normal) > kBackfaceTol) {
            out.status[ray] = kError; out.err_kind[ray] = kErrBackface; out.err_patch[ray] = best_obj;
            out.bounces[ray] = bounces; save(); return;
        }
        if (bounces < K) {
            out.hit_patch[ray * K + bounces] = best_obj;
            out.iters[ray * K + bounces] = (uint8_t)h.iters;
            out.hit_xyz[(ray * K + bounces) * 3 + 0] = h.point.x;
            out.hit_xyz[(ray * K + bounces) * 3 + 1] = h.point.y;
            out.hit_xyz[(ray * K + bounces) * 3 + 2] = h.point.z;
        }
        bounces++;
        const double dn = d.x * h.normal.x + d.y * h.normal.y + d.z * h.normal.z;
        d = { d.x - 2.0 * dn * h.normal.x, d.y - 2.0 * dn * h.normal.y, d.z - 2.0 * dn * h.normal.z };
        const double eps = smax(kSelfEpsFactor * S.diag, kSagittaClearance * S.patches[best_obj].max_sag);
        o = { h.point.x + d.x * eps, h.point.y + d.y * eps, h.point.z + d.z * eps };
    }
    out.status[ray] = kTrapped; out.bounces[ray] = bounces; save();
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (ctypes) -- test infrastructure.
// ---------------------------------------------------------------------------
extern "C" {

struct orc_patch_set {
    int64_t n_patches;
    const int32_t* imax; const int32_t* jmax;
    const int64_t* u_off; const int64_t* v_off; const int64_t* f_off;
    const double* unodes; const double* vnodes; const double* fields;
};

int orc_mt(const double* o, const double* d, const double* v0, const double* v1, const double* v2, double eps,
           double* out /* t, point[3], normal[3] */) {
    double t; Vec3 p, n;
    bool h = moller_trumbore({ o[0], o[1], o[2] }, { d[0], d[1], d[2] }, { v0[0], v0[1], v0[2] },
                             { v1[0], v1[1], v1[2] }, { v2[0], v2[1], v2[2] }, eps, t, &p, &n);
    if (h) { out[0] = t; out[1] = p.x; out[2] = p.y; out[3] = p.z; out[4] = n.x; out[5] = n.y; out[6] = n.z; }
    return h ? 1 : 0;
}

int orc_ray_box(const double* o, const double* inv, const double* box, double best_t, double* tmin_out) {
    AABB b; b.min = { box[0], box[1], box[2] }; b.max = { box[3], box[4], box[5] };
    double tm = std::numeric_limits<double>::quiet_NaN();
    bool r = ray_box({ o[0], o[1], o[2] }, inv, b, best_t, tm);
    *tmin_out = tm;
    return r ? 1 : 0;
}

void orc_bary_weights(const double* nodes, int n, double* out) {
    auto w = barycentric_weights(nodes, n);
    std::memcpy(out, w.data(), sizeof(double) * n);
}

void orc_interp2d(const double* xn, int N, const double* yn, int M, const double* F, const double* x, int P,
                  const double* y, int Q, double* out) {
    auto xw = barycentric_weights(xn, N); auto yw = barycentric_weights(yn, M);
    interp2d(xn, N, yn, M, xw.data(), yw.data(), F, x, P, y, Q, out);
}

void orc_basis1d(const double* nodes, int n, double x, double* out) {
    auto w = barycentric_weights(nodes, n);
    basis1d(nodes, w.data(), n, x, out);
}

double orc_eval_field(const double* lu, const double* lv, const double* f, int imax, int jmax) {
    return eval_field(lu, lv, f, imax, jmax);
}

int orc_solve3x3(const double* J, const double* b, double* x) {
    const double JJ[3][3] = { { J[0], J[1], J[2] }, { J[3], J[4], J[5] }, { J[6], J[7], J[8] } };
    return solve3x3(JJ, { b[0], b[1], b[2] }, x) ? 1 : 0;
}

// Builds the BVH over generic objects (boxes [n][6], centroids [n][3]) and
// writes it in the reference's depth-first layout.  Returns the node count.
int64_t orc_bvh_build(const double* boxes, const double* cents, int64_t n, int leaf_threshold, double* node_box,
                      int32_t* node_right, int32_t* node_start, int32_t* node_count, uint8_t* node_axis,
                      int32_t* leaf_objs) {
    std::vector<AABB> B(n); std::vector<Vec3> Cn(n);
    for (int64_t i = 0; i < n; ++i) {
        B[i].min = { boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2] };
        B[i].max = { boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5] };
        Cn[i] = { cents[3 * i], cents[3 * i + 1], cents[3 * i + 2] };
    }
    BVHBuild bb{ B, Cn, leaf_threshold, {}, {} };
    bb.nodes.reserve(2 * n); bb.leaf.reserve(n);
    std::vector<int32_t> objs(n);
    for (int64_t i = 0; i < n; ++i) objs[i] = (int32_t)i;
    bb.build(std::move(objs));
    for (size_t k = 0; k < bb.nodes.size(); ++k) {
        const Node& nd = bb.nodes[k];
        double* b = node_box + 6 * k;
        b[0] = nd.box.min.x; b[1] = nd.box.min.y; b[2] = nd.box.min.z;
        b[3] = nd.box.max.x; b[4] = nd.box.max.y; b[5] = nd.box.max.z;
        node_right[k] = nd.right; node_start[k] = nd.start; node_count[k] = nd.count; node_axis[k] = nd.axis;
    }
    for (size_t k = 0; k < bb.leaf.size(); ++k) leaf_objs[k] = bb.leaf[k];
    return (int64_t)bb.nodes.size();
}

// Visits a BVH (built as above) with objects whose intersect() returns the
// given t (NaN = miss).  Writes the object visit order (up to max_visits) and
// returns the number of visits; *best_obj gets the winning object or -1.
int64_t orc_bvh_visit(const double* boxes, const double* cents, int64_t n, int leaf_threshold,
                      const double* obj_t, const double* o, const double* d, int32_t* visits, int64_t max_visits,
                      int32_t* best_obj_out, double* best_t_out) {
    std::vector<AABB> B(n); std::vector<Vec3> Cn(n);
    for (int64_t i = 0; i < n; ++i) {
        B[i].min = { boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2] };
        B[i].max = { boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5] };
        Cn[i] = { cents[3 * i], cents[3 * i + 1], cents[3 * i + 2] };
    }
    BVHBuild bb{ B, Cn, leaf_threshold, {}, {} };
    std::vector<int32_t> objs(n);
    for (int64_t i = 0; i < n; ++i) objs[i] = (int32_t)i;
    bb.build(std::move(objs));
    int64_t nv = 0; Counters C; int32_t best, err;
    PatchHit h = bvh_intersect(bb.nodes, bb.leaf, { o[0], o[1], o[2] }, { d[0], d[1], d[2] },
        [&](int32_t k) {
            if (nv < max_visits) visits[nv] = k;
            ++nv;
            PatchHit r{ 0, 0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
            if (!std::isnan(obj_t[k])) { r.status = 1; r.t = obj_t[k]; }
            return r;
        }, best, err, C);
    *best_obj_out = best; *best_t_out = (best >= 0) ? h.t : std::numeric_limits<double>::quiet_NaN();
    return nv;
}

void* orc_scene_create(const orc_patch_set* ps, int threads) {
    auto* S = new Scene();
    const int64_t n = ps->n_patches;
    S->patches.resize(n);
    if (threads > 0) omp_set_num_threads(threads);
    #pragma omp parallel for schedule(dynamic)
    for (int64_t k = 0; k < n; ++k) {
        Patch& P = S->patches[k];
        P.imax = ps->imax[k]; P.jmax = ps->jmax[k];
        P.un.assign(ps->unodes + ps->u_off[k], ps->unodes + ps->u_off[k + 1]);
        P.vn.assign(ps->vnodes + ps->v_off[k], ps->vnodes + ps->v_off[k + 1]);
        P.uw = barycentric_weights(P.un.data(), P.imax);   // SurfacePatch.hpp:335-336
        P.vw = barycentric_weights(P.vn.data(), P.jmax);
        const size_t nn = (size_t)P.imax * P.jmax;
        const double* base = ps->fields + 12 * ps->f_off[k];
        for (int f = 0; f < 12; ++f) P.f[f].assign(base + f * nn, base + (f + 1) * nn);
        P.centroid_and_box();
        P.coarse_grid();
        P.cell_padding();
    }
    // RayTracingDriver.hpp:375-384 -> BVH(objects) with the default threshold
    std::vector<AABB> B(n); std::vector<Vec3> Cn(n);
    for (int64_t k = 0; k < n; ++k) { B[k] = S->patches[k].box; Cn[k] = S->patches[k].centroid; }
    BVHBuild bb{ B, Cn, kLeafThreshold, {}, {} };
    bb.nodes.reserve(2 * n); bb.leaf.reserve(n);
    std::vector<int32_t> objs(n);
    for (int64_t i = 0; i < n; ++i) objs[i] = (int32_t)i;
    bb.build(std::move(objs));
    S->nodes = std::move(bb.nodes); S->leaf = std::move(bb.leaf);
    // RayTracingDriver.hpp:386-396
    AABB box = S->patches.front().box;
    for (const auto& P : S->patches) { box = union_box(box, P.box); S->max_patch_sag = smax(S->max_patch_sag, P.max_sag); }
    const double dx = box.max.x - box.min.x, dy = box.max.y - box.min.y, dz = box.max.z - box.min.z;
    S->diag = std::sqrt(dx * dx + dy * dy + dz * dz);
    return S;
}

void orc_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

double orc_scene_diag(void* s) { return static_cast<Scene*>(s)->diag; }

int64_t orc_scene_n_nodes(void* s) { return (int64_t) static_cast<Scene*>(s)->nodes.size(); }

void orc_scene_bvh(void* s, double* node_box, int32_t* node_right, int32_t* node_start, int32_t* node_count,
                   uint8_t* node_axis, int32_t* leaf_objs) {
    const Scene* S = static_cast<Scene*>(s);
    for (size_t k = 0; k < S->nodes.size(); ++k) {
        const Node& nd = S->nodes[k];
        double* b = node_box + 6 * k;
        b[0] = nd.box.min.x; b[1] = nd.box.min.y; b[2] = nd.box.min.z;
        b[3] = nd.box.max.x; b[4] = nd.box.max.y; b[5] = nd.box.max.z;
        node_right[k] = nd.right; node_start[k] = nd.start; node_count[k] = nd.count; node_axis[k] = nd.axis;
    }
    for (size_t k = 0; k < S->leaf.size(); ++k) leaf_objs[k] = S->leaf[k];
}

// Per patch: box [6], centroid [3], max_sagitta.
void orc_scene_patch_info(void* s, double* box, double* cent, double* max_sag) {
    const Scene* S = static_cast<Scene*>(s);
    for (size_t k = 0; k < S->patches.size(); ++k) {
        const Patch& P = S->patches[k];
        box[6 * k] = P.box.min.x; box[6 * k + 1] = P.box.min.y; box[6 * k + 2] = P.box.min.z;
        box[6 * k + 3] = P.box.max.x; box[6 * k + 4] = P.box.max.y; box[6 * k + 5] = P.box.max.z;
        cent[3 * k] = P.centroid.x; cent[3 * k + 1] = P.centroid.y; cent[3 * k + 2] = P.centroid.z;
        max_sag[k] = P.max_sag;
    }
}

// Coarse grid of one patch: xyz [100*100][3], sagitta [99*99], cell normal [99*99][3].
void orc_scene_coarse(void* s, int64_t k, double* xyz, double* sag, double* cnorm) {
    const Patch& P = static_cast<Scene*>(s)->patches[k];
    for (size_t i = 0; i < P.cx.size(); ++i) { xyz[3 * i] = P.cx[i]; xyz[3 * i + 1] = P.cy[i]; xyz[3 * i + 2] = P.cz[i]; }
    for (size_t i = 0; i < P.sag.size(); ++i) {
        sag[i] = P.sag[i]; cnorm[3 * i] = P.cnorm[i].x; cnorm[3 * i + 1] = P.cnorm[i].y; cnorm[3 * i + 2] = P.cnorm[i].z;
    }
}

// One ray against one patch (SurfacePatch::intersect).  out = t, point[3], normal[3]; returns
// 1 hit, 0 miss, -kind on an error; *iters the Newton iteration count.
int orc_patch_intersect(void* s, int64_t k, const double* o, const double* d, int mode, double* out, int* iters) {
    Counters C;
    PatchHit h = patch_intersect(static_cast<Scene*>(s)->patches[k], { o[0], o[1], o[2] }, { d[0], d[1], d[2] },
                                 mode, C);
    out[0] = h.t; out[1] = h.point.x; out[2] = h.point.y; out[3] = h.point.z;
    out[4] = h.normal.x; out[5] = h.normal.y; out[6] = h.normal.z;
    *iters = h.iters;
    return h.status;
}

// orc_patch_intersect with the converged parameters: out[7] = u, out[8] = v.
int orc_patch_intersect_uv(void* s, int64_t k, const double* o, const double* d, int mode, double* out, int* iters) {
    Counters C;
    PatchHit h = patch_intersect(static_cast<Scene*>(s)->patches[k], { o[0], o[1], o[2] }, { d[0], d[1], d[2] },
                                 mode, C);
    out[0] = h.t; out[1] = h.point.x; out[2] = h.point.y; out[3] = h.point.z;
    out[4] = h.normal.x; out[5] = h.normal.y; out[6] = h.normal.z; out[7] = h.u; out[8] = h.v;
    *iters = h.iters;
    return h.status;
}

// Stage 1 alone and the Newton seed it gives (SurfacePatch.hpp:272-302): returns 1 on a
// stage-1 hit with out = t0, u0, v0 and idx = i, j, second triangle; 0 on a miss.
int orc_stage1_seed(void* s, int64_t k, const double* o, const double* d, double* out, int* idx) {
    Counters C;
    const Patch& P = static_cast<Scene*>(s)->patches[k];
    double bt = 0; int bi = -1, bj = -1; bool bsec = false;
    if (!stage1(P, { o[0], o[1], o[2] }, { d[0], d[1], d[2] }, bt, bi, bj, bsec, C)) return 0;
    double u0, v0;
    triangle_uv_centroid(P, bi, bj, bsec, u0, v0);
    out[0] = bt; out[1] = u0; out[2] = v0;
    idx[0] = bi; idx[1] = bj; idx[2] = bsec ? 1 : 0;
    return 1;
}

// BVH::intersect (BVH.hpp:75-128) over the scene's patches for one ray, direction as given,
// or with brute brute_force_intersect (BruteForce.hpp:7-15) over the patches in order.
// out = t, point[3], normal[3]; *patch the closest hit's patch (-1 none); returns 0 or
// the error kind of the intersection that threw, with *err_patch.
int orc_scene_intersect(void* s, const double* o, const double* d, int mode, int brute, double* out, int* patch,
                        int* err_patch) {
    const Scene& S = *static_cast<Scene*>(s);
    Counters C;
    const Vec3 O = { o[0], o[1], o[2] }, D = { d[0], d[1], d[2] };
    PatchHit best{ 0, 0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    int32_t best_obj = -1, err_obj = -1;
    int err = 0;
    if (brute) {
        for (size_t k = 0; k < S.patches.size(); ++k) {
            const PatchHit h = patch_intersect(S.patches[k], O, D, mode, C);
            if (h.status < 0) { err = -h.status; err_obj = (int32_t)k; break; }
            if (h.status == 1 && (best_obj < 0 || h.t < best.t)) { best = h; best_obj = (int32_t)k; }
        }
    } else {
        const PatchHit h = bvh_intersect(S.nodes, S.leaf, O, D,
                                         [&](int32_t k) { return patch_intersect(S.patches[k], O, D, mode, C); },
                                         best_obj, err_obj, C);
        if (h.status < 0) err = -h.status; else best = h;
    }
    if (err) best_obj = -1;
    out[0] = best.t; out[1] = best.point.x; out[2] = best.point.y; out[3] = best.point.z;
    out[4] = best.normal.x; out[5] = best.normal.y; out[6] = best.normal.z;
    *patch = best_obj;
    *err_patch = err ? err_obj : -1;
    return err;
}

// Newton alone from a given seed (SurfacePatch::newton_refine).
int orc_newton(void* s, int64_t k, const double* o, const double* d, double u0, double v0, double t0, int mode,
               double* out, int* iters) {
    Counters C;
    PatchHit h = newton_refine(static_cast<Scene*>(s)->patches[k], { o[0], o[1], o[2] }, { d[0], d[1], d[2] },
                               u0, v0, t0, mode, C);
    out[0] = h.t; out[1] = h.point.x; out[2] = h.point.y; out[3] = h.point.z;
    out[4] = h.normal.x; out[5] = h.normal.y; out[6] = h.normal.z;
    *iters = h.iters;
    return h.status;
}

// counters: [newton_calls, newton_iters, candidates, mt_tests, box_tests, polish_fail]
void orc_trace(void* s, const double* origins, const double* dirs, int64_t n, int mode, int max_bounces,
               int record_k, int threads, int8_t* status, int8_t* err_kind, int32_t* bounces, int32_t* err_patch,
               int32_t* hit_patch, double* hit_xyz, uint8_t* iters, double* final_od, int64_t* counters) {
    const Scene& S = *static_cast<Scene*>(s);
    RayOut out{ status, err_kind, bounces, err_patch, hit_patch, hit_xyz, iters, final_od, record_k };
    if (threads > 0) omp_set_num_threads(threads);
    int64_t acc[6] = { 0, 0, 0, 0, 0, 0 };
    #pragma omp parallel
    {
        Counters C;
        #pragma omp for schedule(dynamic)
        for (int64_t r = 0; r < n; ++r)
            trace_ray(S, r, { origins[3 * r], origins[3 * r + 1], origins[3 * r + 2] },
                      { dirs[3 * r], dirs[3 * r + 1], dirs[3 * r + 2] }, max_bounces, mode, out, C);
        #pragma omp critical
        {
            acc[0] += C.newton_calls; acc[1] += C.newton_iters; acc[2] += C.candidates;
            acc[3] += C.mt_tests; acc[4] += C.box_tests; acc[5] += C.polish_fail;
        }
    }
    for (int i = 0; i < 6; ++i) counters[i] = acc[i];
}

int orc_max_threads(void) { return omp_get_max_threads(); }

// BounceMapDriver::run's launch generation (BounceMapDriver.hpp:115-220) for the
// scene patches `ids` in order: SurfacePatch::sample_grid (SurfacePatch.hpp:199-242),
// build_tangent_frame (BounceMapDriver.hpp:499-519), the launch nudge along the
// normal by self_intersect_epsilon(patch) (RayTracingDriver.hpp:260-265) and the
// theta x phi cone.  Writes [n_ids][n*n][theta_n*phi_n][3] origins and directions.
void orc_bounce_map_launches(void* s, const int64_t* ids, int64_t n_ids, int n, int theta_n, int phi_n,
                             double cone_deg, double* origins, double* dirs) {
    const Scene& S = *static_cast<Scene*>(s);
    constexpr double kTwoPi = 6.283185307179586476925286766559;
    const double half = cone_deg / 2.0;
    const double cos_theta_max = std::cos(half * kTwoPi / 360.0);
    std::vector<double> g(n);
    for (int i = 0; i < n; ++i) g[i] = (n == 1) ? 0.0 : -1.0 + 2.0 * static_cast<double>(i) / (n - 1);
    size_t at = 0;
    for (int64_t q = 0; q < n_ids; ++q) {
        const Patch& P = S.patches[ids[q]];
        const double eps = smax(kSelfEpsFactor * S.diag, kSagittaClearance * P.max_sag);
        const size_t nn = (size_t)n * n;
        std::vector<double> F[9];
        const int fidx[9] = { 0, 1, 2, 9, 10, 11, 3, 4, 5 };   // x y z, nu, dX/du
        for (int f = 0; f < 9; ++f) {
            F[f].resize(nn);
            interp2d(P.un.data(), P.imax, P.vn.data(), P.jmax, P.uw.data(), P.vw.data(), P.f[fidx[f]].data(), g.data(),
                     n, g.data(), n, F[f].data());
        }
        for (size_t k = 0; k < nn; ++k) {
            const Vec3 nrm0 = vnormalize({ F[3][k], F[4][k], F[5][k] });   // sample_grid renormalises
            const Vec3 point = { F[0][k], F[1][k], F[2][k] };
            const Vec3 normal = vnormalize(nrm0);                          // and run() normalises again
            const Vec3 du = { F[6][k], F[7][k], F[8][k] };
            const double du_len2 = vdot(du, du);
            const double du_dot_n = vdot(du, normal);
            const Vec3 du_perp = { du.x - du_dot_n * normal.x, du.y - du_dot_n * normal.y, du.z - du_dot_n * normal.z };
            const double du_perp_len2 = vdot(du_perp, du_perp);
            Vec3 t1;
            if (du_perp_len2 > 1e-6 * 1e-6 * du_len2) t1 = vnormalize(du_perp);
            else {
                const Vec3 a = (std::fabs(normal.x) <= std::fabs(normal.y) && std::fabs(normal.x) <= std::fabs(normal.z))
                                   ? Vec3{ 1.0, 0.0, 0.0 }
                                   : (std::fabs(normal.y) <= std::fabs(normal.z) ? Vec3{ 0.0, 1.0, 0.0 } : Vec3{ 0.0, 0.0, 1.0 });
                t1 = vnormalize(vcross(normal, a));
            }
            const Vec3 t2 = vcross(normal, t1);
            const Vec3 origin = { point.x + normal.x * eps, point.y + normal.y * eps, point.z + normal.z * eps };
            for (int ti = 0; ti < theta_n; ++ti) {
                const double mu = cos_theta_max + (static_cast<double>(ti) + 0.5) / static_cast<double>(theta_n) * (1.0 - cos_theta_max);
                const double sin_theta = std::sqrt(smax(0.0, 1.0 - mu * mu));
                for (int pj = 0; pj < phi_n; ++pj) {
                    const double phi = (static_cast<double>(pj) + 0.5) * kTwoPi / static_cast<double>(phi_n);
                    const double cp = std::cos(phi), sp = std::sin(phi);
                    origins[3 * at] = origin.x; origins[3 * at + 1] = origin.y; origins[3 * at + 2] = origin.z;
                    dirs[3 * at] = t1.x * sin_theta * cp + t2.x * sin_theta * sp + normal.x * mu;
                    dirs[3 * at + 1] = t1.y * sin_theta * cp + t2.y * sin_theta * sp + normal.y * mu;
                    dirs[3 * at + 2] = t1.z * sin_theta * cp + t2.z * sin_theta * sp + normal.z * mu;
                    ++at;
                }
            }
        }
    }
}

}  // extern "C"
