// runtime.h -- host-side runtime shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_common.h"

namespace bbrt {
// The thread-local message behind bbrt_last_error().
extern thread_local std::string g_last_error;
}  // namespace bbrt
using bbrt::g_last_error;

namespace {
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


// Everything a captured trace graph depends on; a replay needs all of it unchanged.
struct GraphKey {
    const double* d_o;
    const double* d_d;
    int64_t n;
    const int32_t* ray_scene;
    bbrt_params params;
    Outs outs;
    const void* scratch;
    unsigned grid;
    int block;
    const void* kern;
    bool pad;
};

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
    // bbrt_trace_device_async: one scratch buffer per stream, so traces on
    // distinct streams can be in flight together; the mutex guards this list
    // and the per-ray sub-scene ids
    struct StreamScratch {
        hipStream_t stream;
        void* p;
        size_t bytes;
    };
    std::vector<StreamScratch> async_scratch;
    std::mutex async_mu;
    int32_t* h_count = nullptr;   // pinned
    // BBRT_FLAG_GRAPH: the captured trace and the arguments it was captured with
    hipStream_t gstream = nullptr;
    hipEvent_t ev_dep = nullptr;
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey{};
    int64_t graph_captures = 0;
    // sub-scenes (one for bbrt_scene_create, many for bbrt_scene_create_batch)
    int32_t n_scenes = 1;
    bool compact = false;
    std::vector<bbrt::SceneInfo> scenes;
    bbrt::SceneInfo* d_scenes = nullptr;
    std::vector<int64_t> patch_begin;   // [n_scenes + 1]
    // per-ray sub-scene ids of the last batch trace, rebuilt when the ray ranges change
    int32_t* d_ray_scene = nullptr;
    int64_t ray_scene_cap = 0;
    std::vector<int64_t> ray_scene_key;
    // learned seed: per-patch frames (fp64 [n][16]) and features (fp32 [n][48]),
    // and the attached network's snapshot (bbrt_scene_set_mlp)
    double* mlp_frame = nullptr;
    float* mlp_pfeat = nullptr;
    float* mlp_h1pre = nullptr;
    void* mlp_wfrag = nullptr;
    float* mlp_params = nullptr;
    bool mlp_ready = false;
    hipEvent_t ev[2] = { nullptr, nullptr };
    std::vector<hipEvent_t> kev;

    ~bbrt_scene_s() {
        for (void* p : allocs) (void)hipFree(p);
        if (d_ray_scene) (void)hipFree(d_ray_scene);
        if (scratch) (void)hipFree(scratch);
        for (auto& a : async_scratch) if (a.p) (void)hipFree(a.p);
        if (h_count) (void)hipHostFree(h_count);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (ev_dep) (void)hipEventDestroy(ev_dep);
        if (gstream) (void)hipStreamDestroy(gstream);
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

// The async scratch of `st`, grown on demand.  Growing waits for the stream
// first: its earlier traces may still use the old buffer.
void* async_scratch_for(bbrt_scene s, hipStream_t st, size_t need) {
    for (auto& a : s->async_scratch) {
        if (a.stream != st) continue;
        if (need > a.bytes) {
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            check(hipFree(a.p), "hipFree(scratch)");
            a.p = nullptr;
            a.bytes = 0;
            check(hipMalloc(&a.p, need), "hipMalloc(scratch)");
            a.bytes = need;
        }
        return a.p;
    }
    void* p = nullptr;
    check(hipMalloc(&p, need), "hipMalloc(scratch)");
    s->async_scratch.push_back({ st, p, need });
    return p;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace
