// drivers.hpp -- C++ drop-in for the reference's host interfaces, over the C ABI
// of libbbrt_hip.so (include/bbrt.h).  Header-only; link with -lbbrt_hip.
//
// Same class names, signatures, argument meaning and error behaviour as the
// reference's headers, with the per-ray trace loop replaced by one batched
// GPU call:
//   ScatObject, Vec3, Ray, AABB, Hit         include/ScatObject.hpp:10-89
//   moller_trumbore                           include/Utils/MollerTrumbore.hpp:32-64
//   SurfacePatch (file format, data, intersect,
//     coarse_grid, sample_grid, max_sagitta)  include/ScatObjects/SurfacePatch.hpp:52-371
//   Triangle                                  include/ScatObjects/Triangle.hpp:9-32
//   BVHNode, BVH                              include/BVH.hpp:56-189
//   brute_force_intersect                     include/BruteForce.hpp:7-15
//   RayTracingDriver                          include/RayTracingDriver.hpp:65-425
//   PlaneSourceDriver                         include/PlaneSourceDriver.hpp:55-244
//   BounceMapDriver                           include/BounceMapDriver.hpp:61-520
// There is no CPU intersection path: every trace and every intersect() call
// runs the HIP kernels (moller_trumbore, Triangle::intersect, BVH::intersect
// and SurfacePatch::intersect through bbrt_moller_trumbore,
// bbrt_triangles_intersect and bbrt_scene_intersect).
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../bbrt.h"

namespace bbrt_cpp {

// --- value types (ScatObject.hpp:10-47) ----------------------------------
struct Vec3 { double x, y, z; };
struct Ray { Vec3 origin; Vec3 dir; };
struct AABB {
    Vec3 min{ std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
              std::numeric_limits<double>::infinity() };
    Vec3 max{ -std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity(),
              -std::numeric_limits<double>::infinity() };
};
class ScatObject;
struct Hit {
    double t;
    Vec3 point;
    Vec3 normal;
    const ScatObject* object = nullptr;
};

// ScatObject.hpp:55-89
class ScatObject {
public:
    virtual ~ScatObject() = default;
    const AABB& bounding_box() const { return box_; }
    const Vec3& centroid() const { return centroid_; }
    virtual std::optional<Hit> intersect(const Ray& ray) const = 0;
    virtual double self_intersect_padding() const { return 0.0; }
protected:
    AABB box_;
    Vec3 centroid_{ 0, 0, 0 };
};

inline Vec3 vec_sub(const Vec3& a, const Vec3& b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline Vec3 vec_cross(const Vec3& a, const Vec3& b) {
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
inline double vec_dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 vec_normalize(const Vec3& v) {
    double len = std::sqrt(vec_dot(v, v));
    return { v.x / len, v.y / len, v.z / len };
}

namespace detail {
inline void check(int rc, const char* what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + bbrt_last_error());
}
// std::min/std::max of the reference: a NaN second argument is ignored.
inline double lesser(double a, double b) { return (b < a) ? b : a; }
inline double greater(double a, double b) { return (a < b) ? b : a; }

// PatchInterpolation.hpp:25-52
inline std::vector<double> barycentric_weights(const std::vector<double>& x) {
    const int n = (int)x.size();
    std::vector<double> w(n, 1.0);
    for (int j = 0; j < n; ++j) {
        double p = 1.0;
        for (int i = 0; i < n; ++i)
            if (i != j) p *= (x[j] - x[i]);
        w[j] = 1.0 / p;
    }
    return w;
}

// First-form tensor interpolation (PatchInterpolation.hpp:54-159) with the two
// matrix products summed over their inner index upwards from 0.0.
inline void interpolate_2d(const std::vector<double>& xn, const std::vector<double>& yn, const std::vector<double>& xw,
                           const std::vector<double>& yw, const std::vector<double>& F, const std::vector<double>& x,
                           const std::vector<double>& y, double* out) {
    const int N = (int)xn.size(), M = (int)yn.size(), P = (int)x.size(), Q = (int)y.size();
    auto rows = [](const std::vector<double>& nodes, const std::vector<double>& w, const std::vector<double>& pts,
                   std::vector<double>& r, std::vector<double>& scale) {
        const int n = (int)nodes.size(), m = (int)pts.size();
        r.assign((size_t)m * n, 0.0);
        scale.assign(m, 1.0);
        for (int i = 0; i < m; ++i) {
            double prod = 1.0;
            int hot = -1;
            for (int k = 0; k < n; ++k) {
                const double diff = pts[i] - nodes[k];
                if (std::fabs(diff) < 1.0e-12) { hot = k; break; }
                r[(size_t)i * n + k] = w[k] / diff;
                prod *= diff;
            }
            if (hot >= 0) {
                for (int k = 0; k < n; ++k) r[(size_t)i * n + k] = 0.0;
                r[(size_t)i * n + hot] = 1.0;
                scale[i] = 1.0;
            } else {
                scale[i] = prod;
            }
        }
    };
    std::vector<double> rx, sx, ry, sy;
    rows(xn, xw, x, rx, sx);
    rows(yn, yw, y, ry, sy);
    std::vector<double> mid((size_t)P * M);
    for (int i = 0; i < P; ++i)
        for (int m = 0; m < M; ++m) {
            double acc = 0.0;
            for (int k = 0; k < N; ++k) acc += rx[(size_t)i * N + k] * F[(size_t)k * M + m];
            mid[(size_t)i * M + m] = acc;
        }
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < Q; ++j) {
            double acc = 0.0;
            for (int m = 0; m < M; ++m) acc += mid[(size_t)i * M + m] * ry[(size_t)j * M + m];
            out[(size_t)i * Q + j] = acc * (sx[i] * sy[j]);
        }
}
// The messages of the intersection failures (SurfacePatch.hpp:618-646,
// RayTracingDriver.hpp:176-226), by error kind; zone is the patch's zoneID.
inline std::string patch_error_text(int kind, const std::string& zone) {
    switch (kind) {
        case BBRT_ERR_SINGULAR:
            return "SurfacePatch::intersect: singular Jacobian during Newton refinement (zoneID " + zone +
                   "), likely a grazing ray";
        case BBRT_ERR_NOCONV:
            return "SurfacePatch::intersect: Newton refinement failed to converge in 20 iterations (zoneID " + zone + ")";
        case BBRT_ERR_DOMAIN:
            return "SurfacePatch::intersect: Newton refinement converged outside the patch domain (zoneID " + zone + ")";
        case BBRT_ERR_BACKFACE:
            return "RayTracingDriver::trace_ray: hit back face of patch (dot(dir, normal) > 0) — ray likely "
                   "tunneled through a triangulation gap";
        case BBRT_ERR_ZERO_DIR: return "RayTracingDriver::trace_ray: zero-length direction";
        default: return "RayTracingDriver::trace_ray: BVH traversal stack overflow";
    }
}
inline std::string patch_error_text(int kind, long long zone) { return patch_error_text(kind, std::to_string(zone)); }

}  // namespace detail

// --- SurfacePatch (SurfacePatch.hpp:52-371) --------------------------------
class SurfacePatch : public ScatObject {
public:
    struct Data {
        int imax = 0, jmax = 0, kmax = 0;
        long long zoneID = 0;
        std::vector<double> uNodes, vNodes, uWeights, vWeights;
        std::vector<double> x, y, z;
        std::vector<int> mask;
        std::vector<double> dxdu, dydu, dzdu, dxdv, dydv, dzdv, dS, nuX, nuY, nuZ;
        std::array<int, 4> edge_flags{ { 0, 0, 0, 0 } };
        long long point_count() const { return static_cast<long long>(imax) * jmax * kmax; }
    };
    struct SampledGrid {
        int n = 0;
        std::vector<double> u, v, x, y, z, nx, ny, nz, dxdu, dydu, dzdu;
        Vec3 point(int i, int j) const { const int k = i * n + j; return { x[k], y[k], z[k] }; }
        Vec3 normal(int i, int j) const { const int k = i * n + j; return { nx[k], ny[k], nz[k] }; }
        Vec3 tangent_u(int i, int j) const { const int k = i * n + j; return { dxdu[k], dydu[k], dzdu[k] }; }
    };

    // SurfacePatch.hpp:95-115
    struct CoarseGrid {
        int n = 0;
        std::vector<double> u, v;          // n each: -1 + 2 i / (n - 1)
        std::vector<double> x, y, z;       // n * n, idx = i * n + j
        std::vector<double> cell_sagitta;  // (n - 1)^2
        std::vector<Vec3> cell_normal;     // (n - 1)^2
        Vec3 point(int i, int j) const { const int idx = i * n + j; return { x[idx], y[idx], z[idx] }; }
        int cell_index(int i, int j) const { return i * (n - 1) + j; }
    };
    // A patch a driver constructs: its scene build fills in the inflated box and
    // the sagitta (set_prepared), so the patch is not prepared on its own.
    struct Deferred {};

    // SurfacePatch.hpp:125-131: load, then the coarse grid's cell padding inflates
    // the box and gives max_sagitta (computed by the library's scene preparation).
    explicit SurfacePatch(const std::string& path) : SurfacePatch(path, Deferred{}) { prepare(); }
    // From values already parsed (the binary patch cache).
    explicit SurfacePatch(Data d) : SurfacePatch(std::move(d), Deferred{}) { prepare(); }
    SurfacePatch(const std::string& directory, const std::string& file_prefix, long long index)
        : SurfacePatch(directory + file_prefix + std::to_string(index + 1) + ".txt") {}
    SurfacePatch(const std::string& path, Deferred) : data_(load_patch_file(path)) { compute_centroid_and_box(); }
    SurfacePatch(Data d, Deferred) : data_(std::move(d)) { compute_centroid_and_box(); }

    const Data& data() const { return data_; }
    double max_sagitta() const { return max_sagitta_; }
    double self_intersect_padding() const override { return max_sagitta_; }

    // Newton stopping rule of intersect(): BBRT_NEWTON_REF is the reference's.
    int newton_mode = BBRT_NEWTON_REF;

    // SurfacePatch.hpp:253-303 (stage 1 on the padded coarse triangles, Newton from
    // the closest one's centroid), on the GPU: a one-patch scene of this patch,
    // built on the first call.  Newton failures throw std::runtime_error with the
    // reference's messages (SurfacePatch.hpp:618-646).
    std::optional<Hit> intersect(const Ray& ray) const override {
        const double o[3] = { ray.origin.x, ray.origin.y, ray.origin.z };
        const double d[3] = { ray.dir.x, ray.dir.y, ray.dir.z };
        int32_t patch = -1, err_patch = -1;
        int8_t err_kind = 0;
        double t = 0.0, p[3], nrm[3];
        detail::check(bbrt_scene_intersect(device_scene(), o, d, 1, newton_mode, 1, &patch, &t, p, nrm, &err_kind,
                                           &err_patch), "bbrt_scene_intersect");
        if (err_kind) throw std::runtime_error(detail::patch_error_text(err_kind, data_.zoneID));
        if (patch < 0) return std::nullopt;
        Hit h;
        h.t = t; h.point = { p[0], p[1], p[2] }; h.normal = { nrm[0], nrm[1], nrm[2] };
        return h;
    }

    // SurfacePatch.hpp:142: the 100 x 100 coarse grid and its cell padding, as the
    // library's scene preparation computes them (computed on the first call).
    const CoarseGrid& coarse_grid() const {
        std::call_once(coarse_->once, [this] {
            const int n = 100;
            std::vector<double> xyz((size_t)3 * n * n), cell((size_t)4 * (n - 1) * (n - 1));
            OnePatchSet one(data_);
            detail::check(bbrt_prepare_patches(&one.set, nullptr, nullptr, nullptr, xyz.data(), cell.data()),
                          "bbrt_prepare_patches");
            CoarseGrid& g = coarse_->grid;
            g.n = n;
            g.u.resize(n); g.v.resize(n);
            for (int i = 0; i < n; ++i) g.u[i] = g.v[i] = -1.0 + 2.0 * static_cast<double>(i) / (n - 1);
            g.x.resize((size_t)n * n); g.y.resize((size_t)n * n); g.z.resize((size_t)n * n);
            for (size_t k = 0; k < (size_t)n * n; ++k) { g.x[k] = xyz[3 * k]; g.y[k] = xyz[3 * k + 1]; g.z[k] = xyz[3 * k + 2]; }
            const size_t nc = (size_t)(n - 1) * (n - 1);
            g.cell_sagitta.resize(nc); g.cell_normal.resize(nc);
            for (size_t k = 0; k < nc; ++k) {
                g.cell_sagitta[k] = cell[4 * k];
                g.cell_normal[k] = { cell[4 * k + 1], cell[4 * k + 2], cell[4 * k + 3] };
            }
        });
        return coarse_->grid;
    }

    // SurfacePatch.hpp:199-242
    SampledGrid sample_grid(int n) const {
        SampledGrid g;
        g.n = n;
        g.u.resize(n); g.v.resize(n);
        for (int i = 0; i < n; ++i) {
            const double t = (n == 1) ? 0.0 : -1.0 + 2.0 * static_cast<double>(i) / (n - 1);
            g.u[i] = t; g.v[i] = t;
        }
        const size_t nn = (size_t)n * n;
        std::vector<double>* outs[9] = { &g.x, &g.y, &g.z, &g.nx, &g.ny, &g.nz, &g.dxdu, &g.dydu, &g.dzdu };
        const std::vector<double>* ins[9] = { &data_.x, &data_.y, &data_.z, &data_.nuX, &data_.nuY, &data_.nuZ,
                                              &data_.dxdu, &data_.dydu, &data_.dzdu };
        for (int f = 0; f < 9; ++f) {
            outs[f]->resize(nn);
            detail::interpolate_2d(data_.uNodes, data_.vNodes, data_.uWeights, data_.vWeights, *ins[f], g.u, g.v,
                                   outs[f]->data());
        }
        for (size_t k = 0; k < nn; ++k) {
            const Vec3 nrm = vec_normalize({ g.nx[k], g.ny[k], g.nz[k] });
            g.nx[k] = nrm.x; g.ny[k] = nrm.y; g.nz[k] = nrm.z;
        }
        return g;
    }

    // Filled in by the owning driver from bbrt_scene_patch_info (the inflated box and sagitta).
    void set_prepared(const double* box, double max_sag) {
        box_.min = { box[0], box[1], box[2] };
        box_.max = { box[3], box[4], box[5] };
        max_sagitta_ = max_sag;
    }

    // SurfacePatch.hpp:313-371
    static Data load_patch_file(const std::string& path) {
        std::ifstream fin(path);
        if (!fin.is_open()) throw std::runtime_error("SurfacePatch: cannot open patch file: " + path);
        Data d;
        fin >> d.imax >> d.jmax >> d.kmax >> d.zoneID;
        if (!fin) throw std::runtime_error("SurfacePatch: bad header in patch file: " + path);
        const long long n = d.point_count();
        auto read_vec = [&](std::vector<double>& v, long long count) {
            v.resize(count);
            for (long long i = 0; i < count; ++i) fin >> v[i];
        };
        read_vec(d.uNodes, d.imax);
        read_vec(d.vNodes, d.jmax);
        d.uWeights = detail::barycentric_weights(d.uNodes);
        d.vWeights = detail::barycentric_weights(d.vNodes);
        read_vec(d.x, n); read_vec(d.y, n); read_vec(d.z, n);
        d.mask.resize(n);
        for (long long i = 0; i < n; ++i) fin >> d.mask[i];
        read_vec(d.dxdu, n); read_vec(d.dydu, n); read_vec(d.dzdu, n);
        read_vec(d.dxdv, n); read_vec(d.dydv, n); read_vec(d.dzdv, n);
        read_vec(d.dS, n);
        read_vec(d.nuX, n); read_vec(d.nuY, n); read_vec(d.nuZ, n);
        if (!fin) throw std::runtime_error("SurfacePatch: truncated patch file: " + path);
        fin >> std::ws;
        if (!fin.eof()) {
            if (!(fin >> d.edge_flags[0] >> d.edge_flags[1] >> d.edge_flags[2] >> d.edge_flags[3]))
                throw std::runtime_error("SurfacePatch: malformed sharp-edge data in patch file: " + path);
        }
        return d;
    }

private:
    Data data_;
    double max_sagitta_ = 0.0;
    struct DeviceScene {
        std::once_flag once;
        bbrt_scene scene = nullptr;
        ~DeviceScene() { if (scene) bbrt_scene_destroy(scene); }
    };
    struct CoarseCache {
        std::once_flag once;
        CoarseGrid grid;
    };
    // A bbrt_patch_set of this one patch (SurfacePatch.hpp:74-80 field order).
    struct OnePatchSet {
        int32_t imax, jmax;
        int64_t uo[2], vo[2], fo[2];
        std::vector<double> fields;
        bbrt_patch_set set;
        explicit OnePatchSet(const Data& d) : imax(d.imax), jmax(d.jmax) {
            if (d.kmax != 1) throw std::runtime_error("SurfacePatch: only surface patches (kmax == 1) are supported");
            const size_t nn = (size_t)d.imax * d.jmax;
            uo[0] = vo[0] = fo[0] = 0; uo[1] = d.imax; vo[1] = d.jmax; fo[1] = (int64_t)nn;
            const std::vector<double>* src[12] = { &d.x, &d.y, &d.z, &d.dxdu, &d.dydu, &d.dzdu,
                                                   &d.dxdv, &d.dydv, &d.dzdv, &d.nuX, &d.nuY, &d.nuZ };
            fields.resize(12 * nn);
            for (int f = 0; f < 12; ++f) std::copy(src[f]->begin(), src[f]->begin() + nn, fields.begin() + f * nn);
            set = bbrt_patch_set{ 1, &imax, &jmax, uo, vo, fo, d.uNodes.data(), d.vNodes.data(), fields.data() };
        }
    };
    std::shared_ptr<DeviceScene> dev_ = std::make_shared<DeviceScene>();
    std::shared_ptr<CoarseCache> coarse_ = std::make_shared<CoarseCache>();

    bbrt_scene device_scene() const {
        std::call_once(dev_->once, [this] {
            OnePatchSet one(data_);
            detail::check(bbrt_scene_create(&one.set, &dev_->scene), "bbrt_scene_create");
        });
        return dev_->scene;
    }

    // The inflated box and max sagitta of this patch alone (SurfacePatch.hpp:446-492).
    void prepare() {
        OnePatchSet one(data_);
        double box[6], cent[3], ms = 0.0;
        detail::check(bbrt_prepare_patches(&one.set, box, cent, &ms, nullptr, nullptr), "bbrt_prepare_patches");
        set_prepared(box, ms);
    }

    // SurfacePatch.hpp:380-407 (the box is replaced by the inflated one once the scene is built)
    void compute_centroid_and_box() {
        const long long n = data_.point_count();
        if (n == 0) throw std::runtime_error("SurfacePatch: patch has no points");
        double sx = 0.0, sy = 0.0, sz = 0.0;
        AABB box;
        for (long long i = 0; i < n; ++i) {
            const double px = data_.x[i], py = data_.y[i], pz = data_.z[i];
            sx += px; sy += py; sz += pz;
            box.min.x = detail::lesser(box.min.x, px); box.min.y = detail::lesser(box.min.y, py);
            box.min.z = detail::lesser(box.min.z, pz);
            box.max.x = detail::greater(box.max.x, px); box.max.y = detail::greater(box.max.y, py);
            box.max.z = detail::greater(box.max.z, pz);
        }
        const double inv_n = 1.0 / static_cast<double>(n);
        centroid_ = { sx * inv_n, sy * inv_n, sz * inv_n };
        box_ = box;
    }
};

// --- moller_trumbore (Utils/MollerTrumbore.hpp:32-64), on the GPU -------------
inline std::optional<Hit> moller_trumbore(const Ray& ray, const Vec3& v0, const Vec3& v1, const Vec3& v2,
                                          double eps = 1e-9) {
    const double tri[9] = { v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z };
    const double o[3] = { ray.origin.x, ray.origin.y, ray.origin.z }, d[3] = { ray.dir.x, ray.dir.y, ray.dir.z };
    int8_t hit = 0;
    double t = 0.0, p[3], n[3];
    detail::check(bbrt_moller_trumbore(tri, o, d, 1, eps, &hit, &t, p, n), "bbrt_moller_trumbore");
    if (!hit) return std::nullopt;
    Hit h;
    h.t = t; h.point = { p[0], p[1], p[2] }; h.normal = { n[0], n[1], n[2] };
    return h;
}

// --- Triangle (ScatObjects/Triangle.hpp:9-32) -----------------------------------
class Triangle : public ScatObject {
public:
    Triangle(Vec3 v0, Vec3 v1, Vec3 v2) : v0_(v0), v1_(v1), v2_(v2) {
        box_.min = { std::min({ v0.x, v1.x, v2.x }), std::min({ v0.y, v1.y, v2.y }), std::min({ v0.z, v1.z, v2.z }) };
        box_.max = { std::max({ v0.x, v1.x, v2.x }), std::max({ v0.y, v1.y, v2.y }), std::max({ v0.z, v1.z, v2.z }) };
        centroid_ = { (v0.x + v1.x + v2.x) / 3.0, (v0.y + v1.y + v2.y) / 3.0, (v0.z + v1.z + v2.z) / 3.0 };
    }
    std::optional<Hit> intersect(const Ray& ray) const override { return moller_trumbore(ray, v0_, v1_, v2_); }
    const Vec3& v0() const { return v0_; }
    const Vec3& v1() const { return v1_; }
    const Vec3& v2() const { return v2_; }
private:
    Vec3 v0_, v1_, v2_;
};

// --- brute_force_intersect (BruteForce.hpp:7-15): every object's own intersect() --
inline std::optional<Hit> brute_force_intersect(const Ray& ray, const std::vector<const ScatObject*>& objects) {
    std::optional<Hit> best;
    for (const ScatObject* obj : objects) {
        auto hit = obj->intersect(ray);
        if (hit && (!best || hit->t < best->t)) best = hit;
    }
    return best;
}

// --- BVH (BVH.hpp:56-189) ---------------------------------------------------------
struct alignas(32) BVHNode {
    AABB box;
    int32_t right_child;   // -1 marks a leaf
    int32_t obj_start;     // leaf: offset into leaf_objects()
    int32_t obj_count;     // leaf: object count, 0 for internal nodes
    uint8_t split_axis;
};

// The reference's median-split tree (built by the library exactly as BVH::build
// builds it, libstdc++ nth_element included) over Triangle or SurfacePatch
// objects, intersected on the GPU: the triangles go to a device triangle set
// (bbrt_triangles_intersect), the patches to a one-scene handle built from them
// (bbrt_scene_intersect; its tree is the same one, with the reference's default
// leaf threshold 4).  Other ScatObject types have no device form and are refused.
class BVH {
public:
    explicit BVH(std::vector<const ScatObject*> objects, int leaf_threshold = 4)
        : objects_(std::move(objects)), leaf_threshold_(leaf_threshold) {
        if (objects_.empty()) throw std::invalid_argument("BVH: no objects");
        if (leaf_threshold < 1) throw std::invalid_argument("BVH: leaf_threshold must be at least 1");
        bool all_tri = true, all_patch = true;
        for (const ScatObject* o : objects_) {
            all_tri = all_tri && dynamic_cast<const Triangle*>(o) != nullptr;
            all_patch = all_patch && dynamic_cast<const SurfacePatch*>(o) != nullptr;
        }
        if (!all_tri && !all_patch)
            throw std::invalid_argument("BVH: objects must be all Triangle or all SurfacePatch (the kinds the GPU "
                                        "intersects)");
        const size_t n = objects_.size();
        std::vector<double> box(6 * n), cent(3 * n);
        for (size_t k = 0; k < n; ++k) {
            const AABB& b = objects_[k]->bounding_box();
            const Vec3& c = objects_[k]->centroid();
            const double bb[6] = { b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z };
            std::copy(bb, bb + 6, box.begin() + 6 * k);
            cent[3 * k] = c.x; cent[3 * k + 1] = c.y; cent[3 * k + 2] = c.z;
        }
        std::vector<double> nb(12 * n);
        std::vector<int32_t> right(2 * n), start(2 * n), count(2 * n), leaf(n);
        std::vector<uint8_t> axis(2 * n);
        const int64_t m = bbrt_build_bvh(box.data(), cent.data(), (int64_t)n, leaf_threshold, nb.data(), right.data(),
                                         start.data(), count.data(), axis.data(), leaf.data());
        if (m < 0) throw std::runtime_error(std::string("bbrt_build_bvh: ") + bbrt_last_error());
        nodes_.resize((size_t)m);
        for (int64_t k = 0; k < m; ++k) {
            BVHNode& nd = nodes_[(size_t)k];
            nd.box.min = { nb[6 * k], nb[6 * k + 1], nb[6 * k + 2] };
            nd.box.max = { nb[6 * k + 3], nb[6 * k + 4], nb[6 * k + 5] };
            nd.right_child = right[k]; nd.obj_start = start[k]; nd.obj_count = count[k]; nd.split_axis = axis[k];
        }
        for (int32_t i : leaf) leaf_objects_.push_back(objects_[(size_t)i]);
        if (all_tri) {
            std::vector<double> tri(9 * n);
            for (size_t k = 0; k < n; ++k) {
                const auto* t = static_cast<const Triangle*>(objects_[k]);
                const Vec3 v[3] = { t->v0(), t->v1(), t->v2() };
                for (int q = 0; q < 3; ++q) { tri[9 * k + 3 * q] = v[q].x; tri[9 * k + 3 * q + 1] = v[q].y; tri[9 * k + 3 * q + 2] = v[q].z; }
            }
            detail::check(bbrt_triangles_create(tri.data(), (int64_t)n, leaf_threshold, &tris_), "bbrt_triangles_create");
        } else {
            if (leaf_threshold != 4)
                throw std::invalid_argument("BVH: a tree of SurfacePatch objects uses the reference's leaf threshold 4");
            build_patch_scene();
        }
    }
    ~BVH() {
        if (tris_) bbrt_triangles_destroy(tris_);
        if (scene_) bbrt_scene_destroy(scene_);
    }
    BVH(const BVH&) = delete;
    BVH& operator=(const BVH&) = delete;

    // BVH.hpp:75-128 for one ray.  A SurfacePatch intersection failure throws, as
    // the exception of SurfacePatch::intersect propagates out of the reference's.
    std::optional<Hit> intersect(const Ray& ray) const { return intersect_all({ ray }).front(); }

    // The same for many rays in one GPU call.  With patches, the first ray whose
    // query fails throws.
    std::vector<std::optional<Hit>> intersect_all(const std::vector<Ray>& rays) const {
        const size_t n = rays.size();
        std::vector<double> o(3 * n), d(3 * n), t(n), p(3 * n), nrm(3 * n);
        for (size_t i = 0; i < n; ++i) {
            o[3 * i] = rays[i].origin.x; o[3 * i + 1] = rays[i].origin.y; o[3 * i + 2] = rays[i].origin.z;
            d[3 * i] = rays[i].dir.x; d[3 * i + 1] = rays[i].dir.y; d[3 * i + 2] = rays[i].dir.z;
        }
        std::vector<int32_t> obj(n, -1), err_patch(n, -1);
        std::vector<int8_t> err(n, 0);
        if (tris_)
            detail::check(bbrt_triangles_intersect(tris_, o.data(), d.data(), (int64_t)n, 0, obj.data(), t.data(),
                                                   p.data(), nrm.data()), "bbrt_triangles_intersect");
        else
            detail::check(bbrt_scene_intersect(scene_, o.data(), d.data(), (int64_t)n, newton_mode, 0, obj.data(),
                                               t.data(), p.data(), nrm.data(), err.data(), err_patch.data()),
                          "bbrt_scene_intersect");
        std::vector<std::optional<Hit>> out(n);
        for (size_t i = 0; i < n; ++i) {
            if (err[i]) {
                const auto* sp = static_cast<const SurfacePatch*>(objects_[(size_t)err_patch[i]]);
                throw std::runtime_error(detail::patch_error_text(err[i], sp->data().zoneID));
            }
            if (obj[i] < 0) continue;
            Hit h;
            h.t = t[i]; h.point = { p[3 * i], p[3 * i + 1], p[3 * i + 2] };
            h.normal = { nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2] };
            h.object = objects_[(size_t)obj[i]];
            out[i] = h;
        }
        return out;
    }

    // Newton stopping rule for SurfacePatch objects (BBRT_NEWTON_REF is the reference's).
    int newton_mode = BBRT_NEWTON_REF;
    // The tree in the reference's depth-first layout (node 0 is the root).
    const std::vector<BVHNode>& nodes() const { return nodes_; }
    const std::vector<const ScatObject*>& leaf_objects() const { return leaf_objects_; }

private:
    std::vector<const ScatObject*> objects_;
    int leaf_threshold_;
    std::vector<BVHNode> nodes_;
    std::vector<const ScatObject*> leaf_objects_;
    bbrt_triangles tris_ = nullptr;
    bbrt_scene scene_ = nullptr;

    void build_patch_scene() {
        const size_t n = objects_.size();
        std::vector<int32_t> imax(n), jmax(n);
        std::vector<int64_t> uo(n + 1, 0), vo(n + 1, 0), fo(n + 1, 0);
        for (size_t k = 0; k < n; ++k) {
            const auto& d = static_cast<const SurfacePatch*>(objects_[k])->data();
            if (d.kmax != 1) throw std::runtime_error("BVH: only surface patches (kmax == 1) are supported");
            imax[k] = d.imax; jmax[k] = d.jmax;
            uo[k + 1] = uo[k] + d.imax; vo[k + 1] = vo[k] + d.jmax; fo[k + 1] = fo[k] + (int64_t)d.imax * d.jmax;
        }
        std::vector<double> un, vn, fields((size_t)12 * fo[n]);
        for (size_t k = 0; k < n; ++k) {
            const auto& d = static_cast<const SurfacePatch*>(objects_[k])->data();
            un.insert(un.end(), d.uNodes.begin(), d.uNodes.end());
            vn.insert(vn.end(), d.vNodes.begin(), d.vNodes.end());
            const std::vector<double>* src[12] = { &d.x, &d.y, &d.z, &d.dxdu, &d.dydu, &d.dzdu,
                                                   &d.dxdv, &d.dydv, &d.dzdv, &d.nuX, &d.nuY, &d.nuZ };
            const size_t nn = (size_t)d.imax * d.jmax;
            for (int f = 0; f < 12; ++f)
                std::copy(src[f]->begin(), src[f]->begin() + nn, fields.begin() + 12 * fo[k] + f * nn);
        }
        bbrt_patch_set ps{ (int64_t)n, imax.data(), jmax.data(), uo.data(), vo.data(), fo.data(),
                           un.data(), vn.data(), fields.data() };
        detail::check(bbrt_scene_create(&ps, &scene_), "bbrt_scene_create");
    }
};

// --- RayTracingDriver (RayTracingDriver.hpp:65-425) --------------------------
// --- Binary patch cache -------------------------------------------------------
// Parsing a large IFGF-RP directory as text dominates a driver's start-up.  The
// cache keeps every parsed SurfacePatch::Data of a file list in one binary file,
// keyed by the files' paths, sizes and modification times; any change to the
// list or to a file makes it stale, and it is then rebuilt from the text.  The
// values are stored bit for bit, so a scene built from the cache is the scene
// built from the text.  Off unless patch_cache_directory() is set.
namespace detail {
inline std::string& patch_cache_dir() {
    static std::string dir;
    return dir;
}

struct PatchFileKey {
    std::string path;
    uint64_t size = 0;
    int64_t mtime = 0;
};

inline std::vector<PatchFileKey> patch_file_keys(const std::vector<std::string>& files) {
    namespace fs = std::filesystem;
    std::vector<PatchFileKey> keys;
    for (const auto& f : files) {
        PatchFileKey k;
        k.path = fs::absolute(f).lexically_normal().string();
        k.size = (uint64_t)fs::file_size(f);
        k.mtime = (int64_t)fs::last_write_time(f).time_since_epoch().count();
        keys.push_back(std::move(k));
    }
    return keys;
}

// The cache file of a file list: FNV-1a of the normalised paths, in the cache directory.
inline std::string patch_cache_file(const std::vector<PatchFileKey>& keys) {
    uint64_t h = 1469598103934665603ull;
    for (const auto& k : keys) {
        for (unsigned char ch : k.path) { h ^= ch; h *= 1099511628211ull; }
        h ^= 0xFF; h *= 1099511628211ull;
    }
    char name[40];
    std::snprintf(name, sizeof(name), "bbrt_patches_%016llx.bin", (unsigned long long)h);
    return (std::filesystem::path(patch_cache_dir()) / name).string();
}

constexpr char kPatchCacheMagic[8] = { 'B', 'B', 'R', 'T', 'P', 'C', '1', '\0' };

template <class T>
inline void put_vec(std::ofstream& f, const std::vector<T>& v) {
    const uint64_t n = v.size();
    f.write(reinterpret_cast<const char*>(&n), sizeof(n));
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(n * sizeof(T)));
}
template <class T>
inline bool get_vec(std::ifstream& f, std::vector<T>& v) {
    uint64_t n = 0;
    if (!f.read(reinterpret_cast<char*>(&n), sizeof(n)) || n > (1ull << 40) / sizeof(T)) return false;
    v.resize(n);
    return (bool)f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
}

inline void write_patch_cache(const std::string& file, const std::vector<PatchFileKey>& keys,
                              const std::vector<const SurfacePatch::Data*>& data) {
    const std::string tmp = file + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return;   // a cache that cannot be written is skipped, not an error
        f.write(kPatchCacheMagic, 8);
        const uint64_t n = keys.size();
        f.write(reinterpret_cast<const char*>(&n), sizeof(n));
        for (const auto& k : keys) {
            const uint64_t len = k.path.size();
            f.write(reinterpret_cast<const char*>(&len), sizeof(len));
            f.write(k.path.data(), (std::streamsize)len);
            f.write(reinterpret_cast<const char*>(&k.size), sizeof(k.size));
            f.write(reinterpret_cast<const char*>(&k.mtime), sizeof(k.mtime));
        }
        for (const SurfacePatch::Data* d : data) {
            const int32_t dims[3] = { d->imax, d->jmax, d->kmax };
            f.write(reinterpret_cast<const char*>(dims), sizeof(dims));
            f.write(reinterpret_cast<const char*>(&d->zoneID), sizeof(d->zoneID));
            for (const auto* v : { &d->uNodes, &d->vNodes, &d->uWeights, &d->vWeights, &d->x, &d->y, &d->z })
                put_vec(f, *v);
            put_vec(f, d->mask);
            for (const auto* v : { &d->dxdu, &d->dydu, &d->dzdu, &d->dxdv, &d->dydv, &d->dzdv, &d->dS, &d->nuX,
                                   &d->nuY, &d->nuZ })
                put_vec(f, *v);
            f.write(reinterpret_cast<const char*>(d->edge_flags.data()), sizeof(d->edge_flags));
        }
        if (!f) return;
    }
    std::error_code ec;
    std::filesystem::rename(tmp, file, ec);
}

// True and the parsed patches when the cache file matches the keys exactly.
inline bool read_patch_cache(const std::string& file, const std::vector<PatchFileKey>& keys,
                             std::vector<SurfacePatch::Data>& out) {
    std::ifstream f(file, std::ios::binary);
    if (!f) return false;
    char magic[8];
    if (!f.read(magic, 8) || std::memcmp(magic, kPatchCacheMagic, 8) != 0) return false;
    uint64_t n = 0;
    if (!f.read(reinterpret_cast<char*>(&n), sizeof(n)) || n != keys.size()) return false;
    for (const auto& k : keys) {
        uint64_t len = 0;
        if (!f.read(reinterpret_cast<char*>(&len), sizeof(len)) || len != k.path.size()) return false;
        std::string path(len, '\0');
        uint64_t size = 0;
        int64_t mtime = 0;
        if (!f.read(&path[0], (std::streamsize)len) || path != k.path) return false;
        if (!f.read(reinterpret_cast<char*>(&size), sizeof(size)) || size != k.size) return false;
        if (!f.read(reinterpret_cast<char*>(&mtime), sizeof(mtime)) || mtime != k.mtime) return false;
    }
    std::vector<SurfacePatch::Data> data(n);
    for (auto& d : data) {
        int32_t dims[3];
        if (!f.read(reinterpret_cast<char*>(dims), sizeof(dims))) return false;
        d.imax = dims[0]; d.jmax = dims[1]; d.kmax = dims[2];
        if (!f.read(reinterpret_cast<char*>(&d.zoneID), sizeof(d.zoneID))) return false;
        for (auto* v : { &d.uNodes, &d.vNodes, &d.uWeights, &d.vWeights, &d.x, &d.y, &d.z })
            if (!get_vec(f, *v)) return false;
        if (!get_vec(f, d.mask)) return false;
        for (auto* v : { &d.dxdu, &d.dydu, &d.dzdu, &d.dxdv, &d.dydv, &d.dzdv, &d.dS, &d.nuX, &d.nuY, &d.nuZ })
            if (!get_vec(f, *v)) return false;
        if (!f.read(reinterpret_cast<char*>(d.edge_flags.data()), sizeof(d.edge_flags))) return false;
        if (d.x.size() != (size_t)d.point_count()) return false;
    }
    out = std::move(data);
    return true;
}

// The patches of a file list: from the cache when it is current, else parsed
// from the text (and the cache rewritten).  from_cache tells which.
inline std::vector<std::unique_ptr<SurfacePatch>> load_patch_list(const std::vector<std::string>& files,
                                                                  bool* from_cache = nullptr) {
    std::vector<std::unique_ptr<SurfacePatch>> patches;
    if (from_cache) *from_cache = false;
    bool all_there = true;
    for (const auto& f : files) all_there = all_there && std::filesystem::is_regular_file(f);
    if (patch_cache_dir().empty() || !all_there) {   // (a missing file fails in the parser, as without a cache)
        for (const auto& f : files) patches.push_back(std::make_unique<SurfacePatch>(f, SurfacePatch::Deferred{}));
        return patches;
    }
    const auto keys = patch_file_keys(files);
    const std::string cache = patch_cache_file(keys);
    std::vector<SurfacePatch::Data> data;
    if (read_patch_cache(cache, keys, data)) {
        for (auto& d : data) patches.push_back(std::make_unique<SurfacePatch>(std::move(d), SurfacePatch::Deferred{}));
        if (from_cache) *from_cache = true;
        return patches;
    }
    std::vector<const SurfacePatch::Data*> ptrs;
    for (const auto& f : files) {
        patches.push_back(std::make_unique<SurfacePatch>(f, SurfacePatch::Deferred{}));
        ptrs.push_back(&patches.back()->data());
    }
    write_patch_cache(cache, keys, ptrs);
    return patches;
}
}  // namespace detail

// Directory for the binary patch cache of every driver built afterwards ("" turns it off).
inline void set_patch_cache_directory(const std::string& dir) { detail::patch_cache_dir() = dir; }

class RayTracingDriver {
public:
    struct RayPath {
        enum class Status { Escaped, Trapped, Error };
        std::vector<Vec3> points;
        Status status = Status::Trapped;
        int bounce_count = 0;
        std::string error_message;
        std::vector<const ScatObject*> hit_objects;
    };

    // Newton mode of the traces (BBRT_NEWTON_REF reproduces the reference's
    // stopping rule; BBRT_NEWTON_CONVERGED polishes the hit points).
    int newton_mode = BBRT_NEWTON_REF;

    explicit RayTracingDriver(const std::string& patch_directory) {
        load_patches(patch_directory);
        build_scene();
    }
    RayTracingDriver(const std::string& directory, const std::string& file_prefix, int first_index_1based,
                     int last_index_1based) {
        load_patches_range(directory, file_prefix, first_index_1based, last_index_1based);
        build_scene();
    }
    ~RayTracingDriver() { if (scene_) bbrt_scene_destroy(scene_); }
    RayTracingDriver(const RayTracingDriver&) = delete;
    RayTracingDriver& operator=(const RayTracingDriver&) = delete;

    // RayTracingDriver.hpp:119-136
    void set_rays(std::vector<Vec3> sources, std::vector<Vec3> directions) {
        if (sources.size() != directions.size())
            throw std::invalid_argument(
                "RayTracingDriver::set_rays: sources and directions must be the same length (paired 1:1) — got " +
                std::to_string(sources.size()) + " sources and " + std::to_string(directions.size()) + " directions");
        for (size_t i = 0; i < directions.size(); ++i) {
            const Vec3& d = directions[i];
            if (d.x * d.x + d.y * d.y + d.z * d.z < 1e-300)
                throw std::invalid_argument("RayTracingDriver::set_rays: direction " + std::to_string(i) +
                                            " has zero length");
        }
        sources_ = std::move(sources);
        directions_ = std::move(directions);
    }

    // RayTracingDriver.hpp:143-150: every configured ray, in one GPU batch.
    void run(int max_bounces = 50) { paths_ = trace_batch(sources_, directions_, max_bounces); }

    const std::vector<RayPath>& paths() const { return paths_; }
    const std::vector<std::unique_ptr<SurfacePatch>>& patches() const { return patches_; }

    long long patch_index_of(const ScatObject* object) const {
        auto it = index_by_ptr_.find(object);
        return it != index_by_ptr_.end() ? static_cast<long long>(it->second) : -1;
    }

    // Index of the patch loaded from `path` (same file), or -1.
    long long patch_index_of_path(const std::string& path) const {
        namespace fs = std::filesystem;
        std::error_code ec;
        for (size_t k = 0; k < paths_loaded_.size(); ++k)
            if (fs::equivalent(paths_loaded_[k], path, ec)) return (long long)k;
        return -1;
    }

    // Raw batched trace: status, error kind, bounce count and failing patch per
    // ray; records only when record_k > 0.  counters may be NULL.
    void trace_raw(const std::vector<double>& o, const std::vector<double>& d, int max_bounces, int record_k,
                   std::vector<int8_t>& status, std::vector<int8_t>& kind, std::vector<int32_t>& bounces,
                   std::vector<int32_t>& err_patch, std::vector<int32_t>* hit_patch, std::vector<double>* hit_xyz,
                   std::vector<double>* final_od, int64_t* hit_count = nullptr, bool count_direct_hit = false,
                   int64_t* counters = nullptr) const {
        const size_t n = o.size() / 3;
        status.resize(n); kind.resize(n); bounces.resize(n); err_patch.resize(n);
        bbrt_params p{};
        p.newton_mode = newton_mode;
        p.seed_mode = BBRT_SEED_TRIANGLE;
        p.max_bounces = std::max(0, max_bounces);
        p.record_k = record_k;
        p.flags = count_direct_hit ? BBRT_FLAG_COUNT_DIRECT_HIT : 0u;
        bbrt_outputs q{};
        q.status = status.data(); q.err_kind = kind.data(); q.bounces = bounces.data(); q.err_patch = err_patch.data();
        if (record_k > 0 && hit_patch) { hit_patch->resize(n * record_k); q.hit_patch = hit_patch->data(); }
        if (record_k > 0 && hit_xyz) { hit_xyz->resize(n * record_k * 3); q.hit_xyz = hit_xyz->data(); }
        if (final_od) { final_od->resize(6 * n); q.final_od = final_od->data(); }
        q.hit_count = hit_count;
        detail::check(bbrt_trace(scene_, o.data(), d.data(), (int64_t)n, &p, &q, counters), "bbrt_trace");
    }

    // RayTracingDriver.hpp:172-242 for one ray (a batch of one on the GPU).
    RayPath trace_ray(const Vec3& origin, const Vec3& direction, int max_bounces) const {
        return trace_batch({ origin }, { direction }, max_bounces).front();
    }

    // The batched form the drivers use: ray i is sources[i] -> directions[i].
    // Zero-length directions come back as Error paths (like trace_ray).
    std::vector<RayPath> trace_batch(const std::vector<Vec3>& sources, const std::vector<Vec3>& directions,
                                     int max_bounces, int64_t* hit_count = nullptr, bool count_direct_hit = false,
                                     int64_t* counters = nullptr) const {
        const size_t n = sources.size();
        if (directions.size() != n) throw std::invalid_argument("RayTracingDriver::trace_batch: size mismatch");
        std::vector<RayPath> out(n);
        if (n == 0) return out;
        if (max_bounces < 0) max_bounces = 0;
        const int K = std::max(1, max_bounces);
        std::vector<double> o(3 * n), d(3 * n), hit_xyz, final_od;
        for (size_t i = 0; i < n; ++i) {
            o[3 * i] = sources[i].x; o[3 * i + 1] = sources[i].y; o[3 * i + 2] = sources[i].z;
            d[3 * i] = directions[i].x; d[3 * i + 1] = directions[i].y; d[3 * i + 2] = directions[i].z;
        }
        std::vector<int8_t> status, kind;
        std::vector<int32_t> bounces, err_patch, hit_patch;
        trace_raw(o, d, max_bounces, K, status, kind, bounces, err_patch, &hit_patch, &hit_xyz, &final_od, hit_count,
                  count_direct_hit, counters);
        for (size_t i = 0; i < n; ++i) {
            RayPath& path = out[i];
            path.points.push_back(sources[i]);
            path.bounce_count = bounces[i];
            for (int b = 0; b < bounces[i] && b < K; ++b) {
                const size_t at = i * K + b;
                path.points.push_back({ hit_xyz[3 * at], hit_xyz[3 * at + 1], hit_xyz[3 * at + 2] });
                path.hit_objects.push_back(patches_[hit_patch[at]].get());
            }
            switch (status[i]) {
                case BBRT_ESCAPED: {
                    path.status = RayPath::Status::Escaped;
                    const double* f = &final_od[6 * i];   // RayTracingDriver.hpp:198-203
                    path.points.push_back({ f[0] + f[3] * diag_, f[1] + f[4] * diag_, f[2] + f[5] * diag_ });
                    break;
                }
                case BBRT_TRAPPED: path.status = RayPath::Status::Trapped; break;
                default:
                    path.status = RayPath::Status::Error;
                    path.error_message = error_text(kind[i], err_patch[i]);
            }
        }
        return out;
    }

    // RayTracingDriver.hpp:260-265
    double self_intersect_epsilon(const ScatObject* leaving_object = nullptr) const {
        const double padding = leaving_object ? leaving_object->self_intersect_padding() : max_patch_sagitta_;
        return detail::greater(1e-9 * diag_, 5.0 * padding);
    }

    // RayTracingDriver.hpp:270-311, with 17 significant digits.
    void write_vtk(const std::string& filename) const {
        std::ofstream out(filename);
        if (!out.is_open()) throw std::runtime_error("RayTracingDriver::write_vtk: cannot open output file: " + filename);
        write_paths_vtk(out, paths_, "Billiard-ball ray traces", "RayID");
    }

    static void write_paths_vtk(std::ostream& out, const std::vector<RayPath>& paths, const char* title,
                                const char* id_name) {
        std::vector<Vec3> pts;
        struct Seg { int a, b, id, k; };
        std::vector<Seg> segs;
        for (size_t r = 0; r < paths.size(); ++r) {
            const int base = (int)pts.size();
            for (const Vec3& p : paths[r].points) pts.push_back(p);
            for (size_t k = 0; k + 1 < paths[r].points.size(); ++k)
                segs.push_back({ base + (int)k, base + (int)k + 1, (int)r, (int)k });
        }
        out << std::setprecision(17);
        out << "# vtk DataFile Version 3.0\n" << title << "\nASCII\nDATASET POLYDATA\n";
        out << "POINTS " << pts.size() << " double\n";
        for (const Vec3& p : pts) out << p.x << ' ' << p.y << ' ' << p.z << '\n';
        out << "LINES " << segs.size() << ' ' << segs.size() * 3 << '\n';
        for (const Seg& s : segs) out << "2 " << s.a << ' ' << s.b << '\n';
        out << "CELL_DATA " << segs.size() << '\n';
        out << "SCALARS " << id_name << " int 1\nLOOKUP_TABLE default\n";
        for (const Seg& s : segs) out << s.id << '\n';
        out << "SCALARS BounceIndex int 1\nLOOKUP_TABLE default\n";
        for (const Seg& s : segs) out << s.k << '\n';
    }

    double scene_diagonal() const { return diag_; }
    bbrt_scene scene() const { return scene_; }

private:
    std::vector<std::unique_ptr<SurfacePatch>> patches_;
    std::unordered_map<const ScatObject*, size_t> index_by_ptr_;
    std::vector<std::string> paths_loaded_;
    std::vector<Vec3> sources_, directions_;
    std::vector<RayPath> paths_;
    bbrt_scene scene_ = nullptr;
    double diag_ = 0.0, max_patch_sagitta_ = 0.0;

  public:
    std::string error_text(int kind, int patch) const {
        const std::string zone = (patch >= 0 && patch < (int)patches_.size())
                                     ? std::to_string(patches_[patch]->data().zoneID) : std::string("?");
        return detail::patch_error_text(kind, zone);   // the throw sites' wording
    }
  private:

    // RayTracingDriver.hpp:346-373: every *.txt in the directory, sorted by path.
    void load_patches(const std::string& patch_directory) {
        namespace fs = std::filesystem;
        if (!fs::exists(patch_directory) || !fs::is_directory(patch_directory))
            throw std::runtime_error("RayTracingDriver: not a directory: " + patch_directory);
        std::vector<fs::path> files;
        for (const auto& e : fs::directory_iterator(patch_directory))
            if (e.is_regular_file() && e.path().extension() == ".txt") files.push_back(e.path());
        std::sort(files.begin(), files.end());
        for (const auto& p : files) paths_loaded_.push_back(p.string());
        patches_ = detail::load_patch_list(paths_loaded_);
        if (patches_.empty())
            throw std::runtime_error("RayTracingDriver: no *.txt patch files found in: " + patch_directory);
    }

    // RayTracingDriver.hpp:410-424
    void load_patches_range(const std::string& directory, const std::string& prefix, int first, int last) {
        if (last < first)
            throw std::invalid_argument("RayTracingDriver: empty patch index range [" + std::to_string(first) + ", " +
                                        std::to_string(last) + "]");
        for (int i = first; i <= last; ++i) paths_loaded_.push_back(directory + prefix + std::to_string(i) + ".txt");
        patches_ = detail::load_patch_list(paths_loaded_);
    }

    void build_scene() {
        const size_t n = patches_.size();
        std::vector<int32_t> imax(n), jmax(n);
        std::vector<int64_t> uo(n + 1, 0), vo(n + 1, 0), fo(n + 1, 0);
        for (size_t k = 0; k < n; ++k) {
            const auto& d = patches_[k]->data();
            if (d.kmax != 1) throw std::runtime_error("RayTracingDriver: only surface patches (kmax == 1) are supported");
            imax[k] = d.imax; jmax[k] = d.jmax;
            uo[k + 1] = uo[k] + d.imax; vo[k + 1] = vo[k] + d.jmax; fo[k + 1] = fo[k] + (int64_t)d.imax * d.jmax;
        }
        std::vector<double> un, vn, fields((size_t)12 * fo[n]);
        for (size_t k = 0; k < n; ++k) {
            const auto& d = patches_[k]->data();
            un.insert(un.end(), d.uNodes.begin(), d.uNodes.end());
            vn.insert(vn.end(), d.vNodes.begin(), d.vNodes.end());
            const std::vector<double>* src[12] = { &d.x, &d.y, &d.z, &d.dxdu, &d.dydu, &d.dzdu,
                                                   &d.dxdv, &d.dydv, &d.dzdv, &d.nuX, &d.nuY, &d.nuZ };
            const size_t nn = (size_t)d.imax * d.jmax;
            for (int f = 0; f < 12; ++f)
                std::copy(src[f]->begin(), src[f]->begin() + nn, fields.begin() + 12 * fo[k] + f * nn);
        }
        bbrt_patch_set ps{ (int64_t)n, imax.data(), jmax.data(), uo.data(), vo.data(), fo.data(),
                           un.data(), vn.data(), fields.data() };
        detail::check(bbrt_scene_create(&ps, &scene_), "bbrt_scene_create");
        std::vector<double> box(6 * n), cent(3 * n), ms(n);
        detail::check(bbrt_scene_patch_info(scene_, box.data(), cent.data(), ms.data()), "bbrt_scene_patch_info");
        for (size_t k = 0; k < n; ++k) {
            patches_[k]->set_prepared(&box[6 * k], ms[k]);
            index_by_ptr_[patches_[k].get()] = k;
        }
        detail::check(bbrt_scene_info(scene_, &diag_, &max_patch_sagitta_, nullptr, nullptr, nullptr),
                      "bbrt_scene_info");
    }
};

// --- PlaneSourceDriver (PlaneSourceDriver.hpp:55-244) -----------------------
class PlaneSourceDriver {
public:
    explicit PlaneSourceDriver(const std::string& patch_directory) : driver_(patch_directory) {}

    // PlaneSourceDriver.hpp:79-109; the per-patch counts accumulate on the device.
    void run(const Vec3& corner0, const Vec3& corner1, const Vec3& direction, int nx, int ny, int max_bounces = 50,
             bool count_direct_hit = false) {
        if (nx < 1 || ny < 1) throw std::invalid_argument("PlaneSourceDriver::run: nx and ny must both be >= 1");
        if (direction.x * direction.x + direction.y * direction.y + direction.z * direction.z < 1e-300)
            throw std::invalid_argument("PlaneSourceDriver::run: direction has zero length");
        const std::vector<Vec3> sources = build_plane_grid(corner0, corner1, nx, ny);
        hit_count_.assign(driver_.patches().size(), 0);
        const std::vector<Vec3> dirs(sources.size(), direction);
        paths_ = driver_.trace_batch(sources, dirs, max_bounces, reinterpret_cast<int64_t*>(hit_count_.data()),
                                     count_direct_hit);
    }

    // PlaneSourceDriver.hpp:115-178 (17 significant digits)
    void write_vtk(const std::string& filename) const {
        if (hit_count_.empty()) throw std::runtime_error("PlaneSourceDriver::write_vtk: no data — call run() first");
        std::ofstream out(filename);
        if (!out.is_open()) throw std::runtime_error("PlaneSourceDriver::write_vtk: cannot open output file: " + filename);
        const auto& patches = driver_.patches();
        std::vector<Vec3> pts;
        std::vector<long long> hits;
        struct Tri { int a, b, c; };
        std::vector<Tri> tris;
        for (size_t p = 0; p < patches.size(); ++p) {
            const auto& d = patches[p]->data();
            const int base = (int)pts.size();
            for (int idx = 0; idx < d.imax * d.jmax; ++idx) { pts.push_back({ d.x[idx], d.y[idx], d.z[idx] }); hits.push_back(hit_count_[p]); }
            for (int i = 0; i + 1 < d.imax; ++i)
                for (int j = 0; j + 1 < d.jmax; ++j) {
                    const int p00 = base + i * d.jmax + j, p10 = base + (i + 1) * d.jmax + j;
                    const int p01 = base + i * d.jmax + j + 1, p11 = base + (i + 1) * d.jmax + j + 1;
                    tris.push_back({ p00, p10, p11 });
                    tris.push_back({ p00, p11, p01 });
                }
        }
        out << std::setprecision(17);
        out << "# vtk DataFile Version 3.0\nPlane-source patch hit counts\nASCII\nDATASET POLYDATA\n";
        out << "POINTS " << pts.size() << " double\n";
        for (const Vec3& p : pts) out << p.x << ' ' << p.y << ' ' << p.z << '\n';
        out << "POLYGONS " << tris.size() << ' ' << tris.size() * 4 << '\n';
        for (const Tri& t : tris) out << "3 " << t.a << ' ' << t.b << ' ' << t.c << '\n';
        out << "POINT_DATA " << pts.size() << '\n' << "SCALARS HitCount long 1\nLOOKUP_TABLE default\n";
        for (long long h : hits) out << h << '\n';
    }

    const RayTracingDriver& driver() const { return driver_; }
    RayTracingDriver& driver() { return driver_; }
    const std::vector<long long>& hit_counts() const { return hit_count_; }
    const std::vector<RayTracingDriver::RayPath>& paths() const { return paths_; }

    // PlaneSourceDriver.hpp:190-243
    static std::vector<Vec3> build_plane_grid(const Vec3& corner0, const Vec3& corner1, int nx, int ny) {
        const double c0[3] = { corner0.x, corner0.y, corner0.z }, c1[3] = { corner1.x, corner1.y, corner1.z };
        const double scale = std::max({ 1.0, std::fabs(c0[0]), std::fabs(c0[1]), std::fabs(c0[2]), std::fabs(c1[0]),
                                        std::fabs(c1[1]), std::fabs(c1[2]) });
        const double tol = 1e-9 * scale;
        int constant_axis = -1, match_count = 0;
        for (int a = 0; a < 3; ++a)
            if (std::fabs(c0[a] - c1[a]) < tol) { constant_axis = a; ++match_count; }
        if (match_count != 1)
            throw std::invalid_argument(
                "PlaneSourceDriver: corner0 and corner1 must agree in exactly one coordinate to define an axis-aligned "
                "plane (got " + std::to_string(match_count) + " matching coordinates)");
        int axis_a = -1, axis_b = -1;
        for (int a = 0; a < 3; ++a) {
            if (a == constant_axis) continue;
            if (axis_a < 0) axis_a = a; else axis_b = a;
        }
        std::vector<double> ga(nx), gb(ny);
        for (int i = 0; i < nx; ++i)
            ga[i] = (nx == 1) ? c0[axis_a] : c0[axis_a] + (c1[axis_a] - c0[axis_a]) * static_cast<double>(i) / (nx - 1);
        for (int j = 0; j < ny; ++j)
            gb[j] = (ny == 1) ? c0[axis_b] : c0[axis_b] + (c1[axis_b] - c0[axis_b]) * static_cast<double>(j) / (ny - 1);
        std::vector<Vec3> pts;
        pts.reserve((size_t)nx * ny);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                double c[3];
                c[constant_axis] = c0[constant_axis];
                c[axis_a] = ga[i];
                c[axis_b] = gb[j];
                pts.push_back({ c[0], c[1], c[2] });
            }
        return pts;
    }

private:
    RayTracingDriver driver_;
    std::vector<long long> hit_count_;
    std::vector<RayTracingDriver::RayPath> paths_;
};

// --- BounceMapDriver (BounceMapDriver.hpp:61-520) ----------------------------
class BounceMapDriver {
public:
    BounceMapDriver(const std::string& patch_directory, const std::string& file_prefix, int first_index_1based,
                    int last_index_1based)
        : driver_(patch_directory), first_index_1based_(first_index_1based) {
        if (last_index_1based < first_index_1based)
            throw std::invalid_argument("BounceMapDriver: empty patch index range [" +
                                        std::to_string(first_index_1based) + ", " + std::to_string(last_index_1based) +
                                        "]");
        for (int i = first_index_1based; i <= last_index_1based; ++i) {
            const std::string path = patch_directory + file_prefix + std::to_string(i) + ".txt";
            auto sp = std::make_unique<SurfacePatch>(path, SurfacePatch::Deferred{});
            // Its launch nudge uses its own max sagitta (BounceMapDriver.hpp:161): the
            // scene's copy of the same file was prepared with exactly that value.
            const long long k = driver_.patch_index_of_path(path);
            if (k < 0) throw std::runtime_error("BounceMapDriver: sampled patch is not in the scene directory: " + path);
            sample_index_.push_back(k);
            const SurfacePatch& twin = *driver_.patches()[k];
            const AABB& b = twin.bounding_box();
            const double box[6] = { b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z };
            sp->set_prepared(box, twin.max_sagitta());
            sample_patches_.push_back(std::move(sp));
        }
    }

    // BounceMapDriver.hpp:103-259 on the GPU (bbrt_bounce_map): the grid points, tangent frames and launch
    // cones are generated on the device (the cos/sin of the phi angles and cos(theta_max) come from the host
    // libm, as in the reference), all launch rays are traced in one persistent launch, and per point the
    // largest bounce count over directions that did not error, the error count and the first direction
    // reaching that count are reduced on the device.  With capture_best_rays, the best direction of each
    // point (first found on ties, :228-235) is traced once more with its full path recorded.
    void run(int position_grid_n, int theta_n, int phi_n, double cone_angle_deg = 180.0, int max_bounces = 50,
             bool capture_best_rays = false) {
        if (position_grid_n < 2) throw std::invalid_argument("BounceMapDriver::run: position_grid_n must be >= 2");
        if (theta_n < 1 || phi_n < 1) throw std::invalid_argument("BounceMapDriver::run: theta_n and phi_n must both be >= 1");
        if (!(cone_angle_deg > 0.0) || cone_angle_deg > 180.0)
            throw std::invalid_argument("BounceMapDriver::run: cone_angle_deg must be in (0, 180]");
        capture_best_rays_ = capture_best_rays;
        const size_t n_patch = sample_patches_.size();
        const size_t pts_per_patch = (size_t)position_grid_n * position_grid_n;
        const size_t n_points = n_patch * pts_per_patch;
        const size_t n_rays = n_points * (size_t)theta_n * phi_n;
        const double half_angle_deg = cone_angle_deg / 2.0;
        std::vector<double> cphi(phi_n), sphi(phi_n), eps(n_patch);
        for (int pj = 0; pj < phi_n; ++pj) {
            const double phi = (static_cast<double>(pj) + 0.5) * kTwoPi / static_cast<double>(phi_n);
            cphi[pj] = std::cos(phi);
            sphi[pj] = std::sin(phi);
        }
        for (size_t p = 0; p < n_patch; ++p) eps[p] = driver_.self_intersect_epsilon(sample_patches_[p].get());
        bbrt_bounce_map_params bp{};
        bp.grid_n = position_grid_n; bp.theta_n = theta_n; bp.phi_n = phi_n;
        bp.max_bounces = std::max(0, max_bounces);
        bp.newton_mode = driver_.newton_mode;
        bp.cos_theta_max = std::cos(half_angle_deg * kTwoPi / 360.0);
        bp.cos_phi = cphi.data(); bp.sin_phi = sphi.data(); bp.eps = eps.data();
        std::vector<double> grid(9 * n_points), best_od(capture_best_rays ? 6 * n_points : 0);
        std::vector<int32_t> maxb(n_points), errc(n_points), bestd(n_points);
        std::vector<double> lo(keep_launch_rays ? 3 * n_rays : 0), ld(keep_launch_rays ? 3 * n_rays : 0);
        bbrt_bounce_map_outputs bo{ grid.data(), maxb.data(), errc.data(), bestd.data(),
                                    capture_best_rays ? best_od.data() : nullptr,
                                    keep_launch_rays ? lo.data() : nullptr, keep_launch_rays ? ld.data() : nullptr };
        detail::check(bbrt_bounce_map(driver_.scene(), sample_index_.data(), (int64_t)n_patch, &bp, &bo, nullptr),
                      "bbrt_bounce_map");
        launches_ = Launches{};
        for (size_t p = 0; p < n_patch; ++p) {
            SurfacePatch::SampledGrid g;
            g.n = position_grid_n;
            g.u.resize(position_grid_n); g.v.resize(position_grid_n);
            for (int i = 0; i < position_grid_n; ++i)
                g.u[i] = g.v[i] = -1.0 + 2.0 * static_cast<double>(i) / (position_grid_n - 1);
            std::vector<double>* f[9] = { &g.x, &g.y, &g.z, &g.nx, &g.ny, &g.nz, &g.dxdu, &g.dydu, &g.dzdu };
            for (int c = 0; c < 9; ++c) {
                f[c]->resize(pts_per_patch);
                for (size_t k = 0; k < pts_per_patch; ++k) (*f[c])[k] = grid[9 * (p * pts_per_patch + k) + c];
            }
            launches_.grids.push_back(std::move(g));
        }
        if (keep_launch_rays) {
            launches_.origins.resize(n_rays); launches_.dirs.resize(n_rays);
            for (size_t r = 0; r < n_rays; ++r) {
                launches_.origins[r] = { lo[3 * r], lo[3 * r + 1], lo[3 * r + 2] };
                launches_.dirs[r] = { ld[3 * r], ld[3 * r + 1], ld[3 * r + 2] };
            }
        }
        patch_max_bounces_.assign(n_patch, std::vector<int>(pts_per_patch, 0));
        patch_error_counts_.assign(n_patch, std::vector<int>(pts_per_patch, 0));
        for (size_t p = 0; p < n_patch; ++p)
            for (size_t k = 0; k < pts_per_patch; ++k) {
                patch_max_bounces_[p][k] = maxb[p * pts_per_patch + k];
                patch_error_counts_[p][k] = errc[p * pts_per_patch + k];
            }
        patch_best_paths_.clear();
        if (capture_best_rays_) {
            std::vector<Vec3> bo_, bd_;
            std::vector<size_t> where;
            for (size_t q = 0; q < n_points; ++q)
                if (bestd[q] >= 0) {
                    bo_.push_back({ best_od[6 * q], best_od[6 * q + 1], best_od[6 * q + 2] });
                    bd_.push_back({ best_od[6 * q + 3], best_od[6 * q + 4], best_od[6 * q + 5] });
                    where.push_back(q);
                }
            const auto paths = driver_.trace_batch(bo_, bd_, max_bounces);
            patch_best_paths_.assign(n_patch, std::vector<RayTracingDriver::RayPath>(pts_per_patch));
            for (size_t i = 0; i < where.size(); ++i)
                patch_best_paths_[where[i] / pts_per_patch][where[i] % pts_per_patch] = paths[i];
        }
    }

    // Keep every launch ray of the last run() in launches() (off by default: production runs have tens of
    // millions of them; the per-point grids are always kept).
    bool keep_launch_rays = false;

    struct Launches {
        std::vector<Vec3> origins, dirs;
        std::vector<SurfacePatch::SampledGrid> grids;
    };

    const std::vector<std::vector<int>>& max_bounces() const { return patch_max_bounces_; }
    const std::vector<std::vector<int>>& error_counts() const { return patch_error_counts_; }
    const std::vector<std::vector<RayTracingDriver::RayPath>>& best_paths() const { return patch_best_paths_; }
    const Launches& launches() const { return launches_; }

    // BounceMapDriver.hpp:267-335 (17 significant digits)
    void write_vtk(const std::string& filename) const {
        if (launches_.grids.empty()) throw std::runtime_error("BounceMapDriver::write_vtk: no data — call run() first");
        std::ofstream out(filename);
        if (!out.is_open()) throw std::runtime_error("BounceMapDriver::write_vtk: cannot open output file: " + filename);
        std::vector<Vec3> pts;
        std::vector<int> mb, ec;
        struct Tri { int a, b, c; };
        std::vector<Tri> tris;
        for (size_t p = 0; p < launches_.grids.size(); ++p) {
            const auto& g = launches_.grids[p];
            const int n = g.n, base = (int)pts.size();
            for (int idx = 0; idx < n * n; ++idx) {
                pts.push_back({ g.x[idx], g.y[idx], g.z[idx] });
                mb.push_back(patch_max_bounces_[p][idx]);
                ec.push_back(patch_error_counts_[p][idx]);
            }
            for (int i = 0; i + 1 < n; ++i)
                for (int j = 0; j + 1 < n; ++j) {
                    const int p00 = base + i * n + j, p10 = base + (i + 1) * n + j, p01 = base + i * n + j + 1,
                              p11 = base + (i + 1) * n + j + 1;
                    tris.push_back({ p00, p10, p11 });
                    tris.push_back({ p00, p11, p01 });
                }
        }
        out << std::setprecision(17);
        out << "# vtk DataFile Version 3.0\nBilliard-ball bounce map\nASCII\nDATASET POLYDATA\n";
        out << "POINTS " << pts.size() << " double\n";
        for (const Vec3& q : pts) out << q.x << ' ' << q.y << ' ' << q.z << '\n';
        out << "POLYGONS " << tris.size() << ' ' << tris.size() * 4 << '\n';
        for (const Tri& t : tris) out << "3 " << t.a << ' ' << t.b << ' ' << t.c << '\n';
        out << "POINT_DATA " << pts.size() << '\n' << "SCALARS MaxBounces int 1\nLOOKUP_TABLE default\n";
        for (int b : mb) out << b << '\n';
        out << "SCALARS ErrorCount int 1\nLOOKUP_TABLE default\n";
        for (int e : ec) out << e << '\n';
    }

    // BounceMapDriver.hpp:346-408
    void write_best_ray_vtks(const std::string& output_directory,
                             const std::string& file_prefix = "best_rays_patch_") const {
        if (!capture_best_rays_)
            throw std::runtime_error("BounceMapDriver::write_best_ray_vtks: best rays weren't captured — call "
                                     "run(..., /*capture_best_rays=*/true) first");
        if (patch_best_paths_.empty())
            throw std::runtime_error("BounceMapDriver::write_best_ray_vtks: no data — call run() first");
        std::filesystem::create_directories(output_directory);
        for (size_t p = 0; p < patch_best_paths_.size(); ++p) {
            const std::string fn = output_directory + "/" + file_prefix +
                                   std::to_string(first_index_1based_ + (int)p) + ".vtk";
            std::ofstream out(fn);
            if (!out.is_open())
                throw std::runtime_error("BounceMapDriver::write_best_ray_vtks: cannot open output file: " + fn);
            RayTracingDriver::write_paths_vtk(out, patch_best_paths_[p], "Best bounce ray per grid point", "PointID");
        }
    }

    const RayTracingDriver& driver() const { return driver_; }

    // BounceMapDriver.hpp:499-519
    static void build_tangent_frame(const Vec3& n, const Vec3& du, Vec3& t1, Vec3& t2) {
        const double du_len2 = vec_dot(du, du);
        const double du_dot_n = vec_dot(du, n);
        const Vec3 du_perp = { du.x - du_dot_n * n.x, du.y - du_dot_n * n.y, du.z - du_dot_n * n.z };
        const double du_perp_len2 = vec_dot(du_perp, du_perp);
        if (du_perp_len2 > kTangentDegenerateRelTol * kTangentDegenerateRelTol * du_len2) {
            t1 = vec_normalize(du_perp);
        } else {
            const Vec3 a = (std::fabs(n.x) <= std::fabs(n.y) && std::fabs(n.x) <= std::fabs(n.z))
                               ? Vec3{ 1.0, 0.0, 0.0 }
                               : (std::fabs(n.y) <= std::fabs(n.z) ? Vec3{ 0.0, 1.0, 0.0 } : Vec3{ 0.0, 0.0, 1.0 });
            t1 = vec_normalize(vec_cross(n, a));
        }
        t2 = vec_cross(n, t1);
    }

private:
    RayTracingDriver driver_;
    std::vector<std::unique_ptr<SurfacePatch>> sample_patches_;
    std::vector<int64_t> sample_index_;   // each sampled patch's index in the driver's scene
    int first_index_1based_ = 1;
    bool capture_best_rays_ = false;
    Launches launches_;
    std::vector<std::vector<int>> patch_max_bounces_, patch_error_counts_;
    std::vector<std::vector<RayTracingDriver::RayPath>> patch_best_paths_;
    static constexpr double kTwoPi = 6.283185307179586476925286766559;
    static constexpr double kTangentDegenerateRelTol = 1e-6;

};

}  // namespace bbrt_cpp
