// bbrt_objects.hip -- the reference's object-level intersection calls on the GPU:
// Moller-Trumbore on (ray, triangle) pairs (Triangle::intersect), a triangle set
// with its BVH (BVH::intersect and brute_force_intersect over Triangles), and
// BVH::intersect / brute_force_intersect over a scene's curved patches
// (SurfacePatch::intersect behind them).  They back the drop-in classes Triangle,
// BVH, brute_force_intersect and SurfacePatch::intersect of include/bbrt/drivers.hpp,
// so a caller that intersects objects directly, like the reference's own test
// tests/Triangle_Intersect.cpp, runs on the same kernels' arithmetic as the
// bounce loop.  One lane per ray; FP64 throughout, no contraction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "runtime.h"

namespace {

// MollerTrumbore.hpp:32-64, with the hit point o + t d and the unit face normal
// normalize(e1 x e2) that moller_trumbore returns (MollerTrumbore.hpp:57-63).
__device__ __forceinline__ bool mt_full(const V3& o, const V3& d, const V3& v0, const V3& v1, const V3& v2, double eps,
                                        double& t_out, V3& p, V3& n) {
    const V3 e1 = sub3(v1, v0);
    const V3 e2 = sub3(v2, v0);
    const V3 pv = cross3(d, e2);
    const double det = dot3(e1, pv);
    if (fabs(det) < eps) return false;
    const double inv = 1.0 / det;
    const V3 tv = sub3(o, v0);
    const double u = dot3(tv, pv) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const V3 qv = cross3(tv, e1);
    const double v = dot3(d, qv) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    const double t = dot3(e2, qv) * inv;
    if (t < eps) return false;
    t_out = t;
    p = { o.x + t * d.x, o.y + t * d.y, o.z + t * d.z };
    const V3 c = cross3(e1, e2);
    const double len = sqrt(dot3(c, c));
    n = { c.x / len, c.y / len, c.z / len };
    return true;
}

__device__ __forceinline__ V3 load3(const double* p) { return { p[0], p[1], p[2] }; }

// One (ray, triangle) pair per lane.
__global__ void k_mt_pairs(const double* __restrict__ tri, const double* __restrict__ org,
                           const double* __restrict__ dir, int64_t n, double eps, int8_t* hit, double* t, double* p,
                           double* nrm) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    double tt = 0.0;
    V3 pp = { 0, 0, 0 }, nn = { 0, 0, 0 };
    const bool h = mt_full(load3(org + 3 * g), load3(dir + 3 * g), load3(tri + 9 * g), load3(tri + 9 * g + 3),
                           load3(tri + 9 * g + 6), eps, tt, pp, nn);
    hit[g] = h ? 1 : 0;
    t[g] = h ? tt : 0.0;
    p[3 * g] = pp.x; p[3 * g + 1] = pp.y; p[3 * g + 2] = pp.z;
    nrm[3 * g] = nn.x; nrm[3 * g + 1] = nn.y; nrm[3 * g + 2] = nn.z;
}

constexpr int kTriStack = 64;   // BVH.hpp:88 (int32_t stack[64])

// BVH::intersect over triangles (BVH.hpp:75-128): the reference's stack walk, FP64
// ray_box with the best-t prune, leaves in order, a hit kept only if t < best_t;
// or brute_force_intersect (BruteForce.hpp:7-15): every triangle in order, kept if
// there is none yet or t is smaller.  obj = -1 for a miss, -2 for a stack overflow.
__global__ void k_tri_walk(const BvhNode* __restrict__ nodes, const int32_t* __restrict__ leaf,
                           const double* __restrict__ tri, int64_t m, const double* __restrict__ org,
                           const double* __restrict__ dir, int64_t n, int brute, int32_t* obj, double* t, double* p,
                           double* nrm) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const V3 o = load3(org + 3 * g), d = load3(dir + 3 * g);
    int32_t best = -1;
    double best_t = INFINITY;
    V3 bp = { 0, 0, 0 }, bn = { 0, 0, 0 };
    if (brute) {
        for (int64_t k = 0; k < m; ++k) {
            double tt;
            V3 pp, nn;
            if (mt_full(o, d, load3(tri + 9 * k), load3(tri + 9 * k + 3), load3(tri + 9 * k + 6), kMtEps, tt, pp, nn) &&
                (best < 0 || tt < best_t)) {
                best = (int32_t)k; best_t = tt; bp = pp; bn = nn;
            }
        }
    } else {
        const double inv[3] = { 1.0 / d.x, 1.0 / d.y, 1.0 / d.z };
        const bool neg[3] = { inv[0] < 0.0, inv[1] < 0.0, inv[2] < 0.0 };
        int32_t stack[kTriStack];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int32_t idx = stack[--sp];
            const BvhNode nd = nodes[idx];
            if (!ray_box64(o, inv, nd, best_t)) continue;
            if (nd.right < 0) {
                for (int32_t i = 0; i < nd.count; ++i) {
                    const int32_t k = leaf[nd.start + i];
                    double tt;
                    V3 pp, nn;
                    if (mt_full(o, d, load3(tri + 9 * (int64_t)k), load3(tri + 9 * (int64_t)k + 3),
                                load3(tri + 9 * (int64_t)k + 6), kMtEps, tt, pp, nn) && tt < best_t) {
                        best = k; best_t = tt; bp = pp; bn = nn;
                    }
                }
                continue;
            }
            if (sp + 2 > kTriStack) { best = -2; break; }
            const int32_t left = idx + 1, right = nd.right;
            if (neg[nd.axis]) { stack[sp++] = left; stack[sp++] = right; }
            else { stack[sp++] = right; stack[sp++] = left; }
        }
    }
    obj[g] = best;
    t[g] = best >= 0 ? best_t : 0.0;
    p[3 * g] = bp.x; p[3 * g + 1] = bp.y; p[3 * g + 2] = bp.z;
    nrm[3 * g] = bn.x; nrm[3 * g + 1] = bn.y; nrm[3 * g + 2] = bn.z;
}

// BVH::intersect over a one-scene handle's patches (BVH.hpp:75-128 with
// SurfacePatch::intersect at the leaves, SurfacePatch.hpp:253-303: the bounce
// kernel's traverse()), or brute_force_intersect over them in patch order.  The
// direction is used as given (BVH::intersect does not normalise it).  FP32
// culling only for rays within the culling limit and directions of moderate
// length, where its widening bound holds; otherwise the exact brute-force
// stage 1.
__global__ void k_scene_isect(DevScene S, int64_t np, const double* __restrict__ org, const double* __restrict__ dir,
                              int64_t n, int mode, int brute, int32_t* patch, double* t, double* p, double* nrm, int8_t* err_kind,
                              int32_t* err_patch, unsigned long long* ctr) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Counters c = {};
    if (g < n) {
        const V3 o = load3(org + 3 * g), d = load3(dir + 3 * g);
        const double dd = dot3(d, d);
        const bool cull = cull_ok(S, 0, o, d, 0u) && dd >= 1e-6 && dd <= 1e6;
        const Ray32 r32 = { (float)o.x, (float)o.y, (float)o.z, 1.0f / (float)d.x, 1.0f / (float)d.y, 1.0f / (float)d.z };
        const double widen = S.scenes[0].cull_widen;
        PHit best{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
        int32_t best_obj = -1, err_obj = -1;
        int err = 0;
        if (brute) {
            for (int64_t k = 0; k < np; ++k) {
                const PHit h = patch_intersect(S, k, o, d, r32, cull, widen, mode, c);
                if (h.status < 0) { err = -h.status; err_obj = (int32_t)k; break; }
                if (h.status == 1 && (best_obj < 0 || h.t < best.t)) { best = h; best_obj = (int32_t)k; }
            }
        } else {
            err = traverse(S, S.scenes[0].root, widen, o, d, r32, cull, mode, false, best, best_obj, err_obj, c);
        }
        if (err) best_obj = -1;
        patch[g] = best_obj;
        err_kind[g] = (int8_t)err;
        err_patch[g] = err ? err_obj : -1;
        t[g] = best_obj >= 0 ? best.t : 0.0;
        p[3 * g] = best_obj >= 0 ? best.p.x : 0.0; p[3 * g + 1] = best_obj >= 0 ? best.p.y : 0.0;
        p[3 * g + 2] = best_obj >= 0 ? best.p.z : 0.0;
        nrm[3 * g] = best_obj >= 0 ? best.n.x : 0.0; nrm[3 * g + 1] = best_obj >= 0 ? best.n.y : 0.0;
        nrm[3 * g + 2] = best_obj >= 0 ? best.n.z : 0.0;
    }
    flush_counters(c, ctr);
}

// Device buffers for one call, freed on scope exit.
struct CallMem {
    std::vector<void*> p;
    ~CallMem() { for (void* q : p) (void)hipFree(q); }
    template <class T>
    T* get(size_t count) {
        void* q = nullptr;
        check(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc(objects)");
        p.push_back(q);
        return static_cast<T*>(q);
    }
    template <class T>
    T* up(const T* src, size_t count) {
        T* d = get<T>(count);
        if (count) check(hipMemcpy(d, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
        return d;
    }
};

template <class T>
void down(T* dst, const T* src, size_t count) {
    if (dst && count) check(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

unsigned blocks_of(int64_t n, int b) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

}  // namespace

struct bbrt_triangles_s {
    int device = 0;
    int64_t m = 0;
    double* tri = nullptr;
    BvhNode* nodes = nullptr;
    int32_t* leaf = nullptr;
    ~bbrt_triangles_s() {
        if (tri) (void)hipFree(tri);
        if (nodes) (void)hipFree(nodes);
        if (leaf) (void)hipFree(leaf);
    }
};

extern "C" {

int bbrt_moller_trumbore(const double* tri, const double* origins, const double* dirs, int64_t n, double eps,
                         int8_t* hit, double* t, double* point, double* normal) {
    if (n < 0 || (n > 0 && (!tri || !origins || !dirs || !hit || !t || !point || !normal))) {
        g_last_error = "bbrt_moller_trumbore: NULL argument or negative count";
        return -1;
    }
    if (n == 0) return 0;
    try {
        CallMem mem;
        const double* d_tri = mem.up(tri, 9 * (size_t)n);
        const double* d_o = mem.up(origins, 3 * (size_t)n);
        const double* d_d = mem.up(dirs, 3 * (size_t)n);
        int8_t* d_hit = mem.get<int8_t>(n);
        double* d_t = mem.get<double>(n);
        double* d_p = mem.get<double>(3 * (size_t)n);
        double* d_n = mem.get<double>(3 * (size_t)n);
        hipLaunchKernelGGL(k_mt_pairs, dim3(blocks_of(n, 256)), dim3(256), 0, 0, d_tri, d_o, d_d, n, eps, d_hit, d_t,
                           d_p, d_n);
        check(hipGetLastError(), "k_mt_pairs");
        down(hit, d_hit, n); down(t, d_t, n); down(point, d_p, 3 * (size_t)n); down(normal, d_n, 3 * (size_t)n);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_triangles_create(const double* tri, int64_t m, int32_t leaf_threshold, bbrt_triangles* out) {
    if (!tri || !out || m <= 0 || m > (int64_t)INT32_MAX || leaf_threshold < 1) {
        g_last_error = "bbrt_triangles_create: bad arguments";
        return -1;
    }
    try {
        // Triangle.hpp:11-23: the exact box of the three vertices and their mean
        std::vector<double> box(6 * (size_t)m), cent(3 * (size_t)m);
        for (int64_t k = 0; k < m; ++k) {
            const double* v = tri + 9 * k;
            for (int a = 0; a < 3; ++a) {
                box[6 * k + a] = std::min({ v[a], v[3 + a], v[6 + a] });
                box[6 * k + 3 + a] = std::max({ v[a], v[3 + a], v[6 + a] });
                cent[3 * k + a] = (v[a] + v[3 + a] + v[6 + a]) / 3.0;
            }
        }
        std::vector<BvhNode> nodes;
        std::vector<int32_t> leaf;
        bbrt::build_bvh(box, cent, leaf_threshold, nodes, leaf);
        auto* s = new bbrt_triangles_s();
        std::unique_ptr<bbrt_triangles_s> hold(s);
        check(hipGetDevice(&s->device), "hipGetDevice");
        s->m = m;
        check(hipMalloc(&s->tri, 9 * (size_t)m * sizeof(double)), "hipMalloc(triangles)");
        check(hipMalloc(&s->nodes, nodes.size() * sizeof(BvhNode)), "hipMalloc(triangles)");
        check(hipMalloc(&s->leaf, leaf.size() * sizeof(int32_t)), "hipMalloc(triangles)");
        check(hipMemcpy(s->tri, tri, 9 * (size_t)m * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(s->nodes, nodes.data(), nodes.size() * sizeof(BvhNode), hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(s->leaf, leaf.data(), leaf.size() * sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy");
        *out = hold.release();
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

void bbrt_triangles_destroy(bbrt_triangles ts) {
    if (!ts) return;
    (void)hipSetDevice(ts->device);
    delete ts;
}

int bbrt_triangles_intersect(bbrt_triangles ts, const double* origins, const double* dirs, int64_t n,
                             int32_t brute_force, int32_t* object, double* t, double* point, double* normal) {
    if (!ts || n < 0 || (n > 0 && (!origins || !dirs || !object || !t || !point || !normal))) {
        g_last_error = "bbrt_triangles_intersect: NULL argument or negative count";
        return -1;
    }
    if (n == 0) return 0;
    try {
        check(hipSetDevice(ts->device), "hipSetDevice");
        CallMem mem;
        const double* d_o = mem.up(origins, 3 * (size_t)n);
        const double* d_d = mem.up(dirs, 3 * (size_t)n);
        int32_t* d_obj = mem.get<int32_t>(n);
        double* d_t = mem.get<double>(n);
        double* d_p = mem.get<double>(3 * (size_t)n);
        double* d_n = mem.get<double>(3 * (size_t)n);
        hipLaunchKernelGGL(k_tri_walk, dim3(blocks_of(n, 64)), dim3(64), 0, 0, ts->nodes, ts->leaf, ts->tri, ts->m,
                           d_o, d_d, n, brute_force ? 1 : 0, d_obj, d_t, d_p, d_n);
        check(hipGetLastError(), "k_tri_walk");
        down(object, d_obj, n); down(t, d_t, n); down(point, d_p, 3 * (size_t)n); down(normal, d_n, 3 * (size_t)n);
        for (int64_t i = 0; i < n; ++i)
            if (object[i] == -2) throw std::runtime_error("bbrt_triangles_intersect: BVH traversal stack overflow");
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_scene_intersect(bbrt_scene s, const double* origins, const double* dirs, int64_t n, int32_t newton_mode,
                         int32_t brute_force, int32_t* patch, double* t, double* point, double* normal,
                         int8_t* err_kind, int32_t* err_patch) {
    if (!s || n < 0 || (n > 0 && (!origins || !dirs || !patch || !t || !point || !normal || !err_kind || !err_patch))) {
        g_last_error = "bbrt_scene_intersect: NULL argument or negative count";
        return -1;
    }
    if (s->n_scenes != 1) { g_last_error = "bbrt_scene_intersect: one-scene handles only"; return -1; }
    if (newton_mode != BBRT_NEWTON_REF && newton_mode != BBRT_NEWTON_CONVERGED) {
        g_last_error = "bbrt_scene_intersect: unknown newton_mode";
        return -1;
    }
    if (n == 0) return 0;
    try {
        check(hipSetDevice(s->device), "hipSetDevice");
        CallMem mem;
        const double* d_o = mem.up(origins, 3 * (size_t)n);
        const double* d_d = mem.up(dirs, 3 * (size_t)n);
        int32_t* d_patch = mem.get<int32_t>(n);
        double* d_t = mem.get<double>(n);
        double* d_p = mem.get<double>(3 * (size_t)n);
        double* d_n = mem.get<double>(3 * (size_t)n);
        int8_t* d_ek = mem.get<int8_t>(n);
        int32_t* d_ep = mem.get<int32_t>(n);
        unsigned long long* d_ctr = mem.get<unsigned long long>(BBRT_NCOUNTERS);
        check(hipMemsetAsync(d_ctr, 0, BBRT_NCOUNTERS * sizeof(unsigned long long), s->stream), "hipMemsetAsync");
        DevScene dev = s->dev;
        dev.ray_scene = nullptr;
        hipLaunchKernelGGL(k_scene_isect, dim3(blocks_of(n, 64)), dim3(64), 0, s->stream, dev, s->n_patches, d_o, d_d, n,
                           newton_mode,
                           brute_force ? 1 : 0, d_patch, d_t, d_p, d_n, d_ek, d_ep, d_ctr);
        check(hipGetLastError(), "k_scene_isect");
        check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        down(patch, d_patch, n); down(t, d_t, n); down(point, d_p, 3 * (size_t)n); down(normal, d_n, 3 * (size_t)n);
        down(err_kind, d_ek, n); down(err_patch, d_ep, n);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

}  // extern "C"
