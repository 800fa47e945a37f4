// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Exposes the reference's own compiled code, through a small C ABI, for the
// parts of the path whose headers build in this image without stand-ins:
// moller_trumbore (include/Utils/MollerTrumbore.hpp:32-64), ray_box
// (include/BVH.hpp:14-37), BVH build and traversal (include/BVH.hpp:65-189),
// brute_force_intersect (include/BruteForce.hpp:7-15) and Triangle
// (include/ScatObjects/Triangle.hpp:9-32).  SurfacePatch.hpp needs <mkl.h>,
// which this image does not have, so it is not built here.
//
// Built by oracle/Makefile into oracle/_ref/libref_harness.so with the
// reference headers read in place from the reference checkout.  Used only to
// write fixtures under tests/golden/ and by CPU tests when it is present.
#include "BVH.hpp"
#include "BruteForce.hpp"
#include "ScatObjects/Triangle.hpp"
#include "Utils/MollerTrumbore.hpp"

#include <cmath>
#include <limits>
#include <vector>

namespace {

// A ScatObject with a given box and centroid whose intersect() records the
// call and returns a fixed t (NaN = miss).
class ProbeObject : public ScatObject {
public:
    ProbeObject(int id, const double* box, const double* cent, double t, std::vector<int>* log)
        : id_(id), t_(t), log_(log) {
        box_.min = { box[0], box[1], box[2] };
        box_.max = { box[3], box[4], box[5] };
        centroid_ = { cent[0], cent[1], cent[2] };
    }
    std::optional<Hit> intersect(const Ray&) const override {
        log_->push_back(id_);
        if (std::isnan(t_)) return std::nullopt;
        Hit h; h.t = t_; h.point = { 0, 0, 0 }; h.normal = { 0, 0, 1 };
        return h;
    }
    int id() const { return id_; }
private:
    int id_;
    double t_;
    std::vector<int>* log_;
};

}  // namespace

extern "C" {

int ref_mt(const double* o, const double* d, const double* v0, const double* v1, const double* v2, double eps,
           double* out) {
    Ray r{ { o[0], o[1], o[2] }, { d[0], d[1], d[2] } };
    auto h = moller_trumbore(r, { v0[0], v0[1], v0[2] }, { v1[0], v1[1], v1[2] }, { v2[0], v2[1], v2[2] }, eps);
    if (!h) return 0;
    out[0] = h->t; out[1] = h->point.x; out[2] = h->point.y; out[3] = h->point.z;
    out[4] = h->normal.x; out[5] = h->normal.y; out[6] = h->normal.z;
    return 1;
}

int ref_ray_box(const double* o, const double* inv, const double* box, double best_t, double* tmin_out) {
    AABB b; b.min = { box[0], box[1], box[2] }; b.max = { box[3], box[4], box[5] };
    double tm = std::numeric_limits<double>::quiet_NaN();
    bool r = ray_box({ o[0], o[1], o[2] }, inv, b, best_t, tm);
    *tmin_out = tm;
    return r ? 1 : 0;
}

// BVH(objects, leaf_threshold).intersect(ray) over probe objects; writes the
// visit order and the winner (index, t).
int64_t ref_bvh_visit(const double* boxes, const double* cents, int64_t n, int leaf_threshold, const double* obj_t,
                      const double* o, const double* d, int32_t* visits, int64_t max_visits, int32_t* best_obj,
                      double* best_t) {
    std::vector<int> log;
    std::vector<ProbeObject> objs;
    objs.reserve(n);
    for (int64_t i = 0; i < n; ++i) objs.emplace_back((int)i, boxes + 6 * i, cents + 3 * i, obj_t[i], &log);
    std::vector<const ScatObject*> ptrs;
    for (auto& p : objs) ptrs.push_back(&p);
    BVH bvh(ptrs, leaf_threshold);
    Ray r{ { o[0], o[1], o[2] }, { d[0], d[1], d[2] } };
    auto h = bvh.intersect(r);
    for (size_t k = 0; k < log.size() && (int64_t)k < max_visits; ++k) visits[k] = log[k];
    if (h) {
        *best_obj = static_cast<const ProbeObject*>(h->object)->id();
        *best_t = h->t;
    } else {
        *best_obj = -1;
        *best_t = std::numeric_limits<double>::quiet_NaN();
    }
    return (int64_t)log.size();
}

// The known-answer scene of tests/Triangle_Intersect.cpp:14-31 through BVH
// and brute force: out[r] = t of the BVH hit (NaN = miss), out[4 + r] = brute t.
void ref_triangle_kat(const double* tri /*[n][9]*/, int n, const double* rays /*[m][6]*/, int m, int leaf_threshold,
                      double* out) {
    std::vector<Triangle> tris;
    for (int i = 0; i < n; ++i) {
        const double* v = tri + 9 * i;
        tris.emplace_back(Vec3{ v[0], v[1], v[2] }, Vec3{ v[3], v[4], v[5] }, Vec3{ v[6], v[7], v[8] });
    }
    std::vector<const ScatObject*> objects;
    for (auto& t : tris) objects.push_back(&t);
    BVH bvh(objects, leaf_threshold);
    for (int r = 0; r < m; ++r) {
        Ray ray{ { rays[6 * r], rays[6 * r + 1], rays[6 * r + 2] }, { rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5] } };
        auto a = bvh.intersect(ray);
        auto b = brute_force_intersect(ray, objects);
        out[r] = a ? a->t : std::numeric_limits<double>::quiet_NaN();
        out[m + r] = b ? b->t : std::numeric_limits<double>::quiet_NaN();
    }
}

}  // extern "C"
