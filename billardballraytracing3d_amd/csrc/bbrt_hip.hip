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
#include <atomic>
#include <thread>
#include <type_traits>
#include <vector>

#include "mlp_device.h"
#include "runtime.h"
#include "bvh_device.h"

thread_local std::string bbrt::g_last_error;

namespace {

// BBRT_FLAG_GROUP_RAYS: a live ray's group bin is the cell of its scene's root
// box (2^bits per axis, Morton order) that holds the ray point closest to the
// box centre, sub-scene major.  The persistent kernel's queue takes the bins in
// order, so the rays a wave takes together pass the same part of the scene.
constexpr int kGroupBitsOne = 5;    // one scene: 32^3 bins
constexpr int kGroupBitsBatch = 2;  // per sub-scene: 4^3 bins
__device__ __forceinline__ int group_bin(const bbrt::BvhNode& rt, const V3& o, const V3& d, int bits, int scn) {
    const double c[3] = { 0.5 * (rt.lo[0] + rt.hi[0]), 0.5 * (rt.lo[1] + rt.hi[1]), 0.5 * (rt.lo[2] + rt.hi[2]) };
    const double s = (c[0] - o.x) * d.x + (c[1] - o.y) * d.y + (c[2] - o.z) * d.z;
    const double t = s > 0.0 ? s : 0.0;
    const double p[3] = { o.x + d.x * t, o.y + d.y * t, o.z + d.z * t };
    const int L = 1 << bits;
    int code = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double f = (p[a] - rt.lo[a]) / (rt.hi[a] - rt.lo[a]) * L;
        const int q = (f >= 0.0 && f < (double)L) ? (int)f : (f >= (double)L ? L - 1 : 0);   // NaN: 0
        for (int b = 0; b < bits; ++b) code |= ((q >> b) & 1) << (3 * b + a);
    }
    return (scn << (3 * bits)) | code;
}

// RayTracingDriver::trace_ray, lines 172-184: zero-direction check and
// normalize_dir (:398-401); then the root box test of the first BVH query
// (BVH.hpp:91-97 with best_t = inf), so rays that miss the whole scene leave
// here as Escaped after one node test.  The rest enter the live list; with
// group_bits > 0 each also gets its group bin (bin_out, at its list position)
// and is counted in hist.
__global__ void __launch_bounds__(1024) k_init(const double* __restrict__ origins, const double* __restrict__ dirs, int64_t n,
                       DevScene S, RayState R, Outs O, int max_bounces, int32_t* live_out,
                       int32_t* n_out, unsigned long long* ctr, int group_bits, int32_t* bin_out, int32_t* hist) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    int32_t r = (int32_t)g;
    int bin = 0;
    unsigned long long root_tests = 0;
    if (g < n) {
        const V3 o = { origins[3 * g], origins[3 * g + 1], origins[3 * g + 2] };
        const V3 dr = { dirs[3 * g], dirs[3 * g + 1], dirs[3 * g + 2] };
        if (O.ray_work) { O.ray_work[4 * g] = 0; O.ray_work[4 * g + 1] = 0; O.ray_work[4 * g + 2] = 0; O.ray_work[4 * g + 3] = 0; }
        if (O.seed_stats) { O.seed_stats[2 * g] = 0; O.seed_stats[2 * g + 1] = 0; }
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
            } else if (!box_static(o, inv, S.bvh[S.scenes[ray_scene_of(S, g)].root], tm)) {
                root_tests = 1;
                if (O.status) O.status[g] = bbrt::kEscaped;
                if (O.bounces) O.bounces[g] = 0;
                if (O.ray_work) O.ray_work[4 * g] = 1;
                write_final(O, r, o, d);
            } else {
                R.o[3 * g] = o.x; R.o[3 * g + 1] = o.y; R.o[3 * g + 2] = o.z;
                R.d[3 * g] = d.x; R.d[3 * g + 1] = d.y; R.d[3 * g + 2] = d.z;
                alive = true;
                if (group_bits > 0) {
                    const int scn = ray_scene_of(S, g);
                    bin = group_bin(S.bvh[S.scenes[scn].root], o, d, group_bits, scn);
                    atomicAdd(hist + bin, 1);
                }
            }
        }
    }
    // One atomic per 1024-thread block for the live list and each counter:
    // per-wave atomics on one address serialise (16k of them for 2^20 rays).
    __shared__ int w_live[16];
    __shared__ unsigned long long w_rays[16], w_tests[16];
    __shared__ int block_base;
    const unsigned long long mask = __ballot(alive);
    const unsigned long long in_range = __ballot(g < n);
    const unsigned long long rt = wave_sum(root_tests);
    const int wid = threadIdx.x >> 6, lane = __lane_id();
    if (lane == 0) { w_live[wid] = (int)__popcll(mask); w_rays[wid] = __popcll(in_range); w_tests[wid] = rt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        int tot = 0;
        unsigned long long rays = 0, tests = 0;
        for (int w = 0; w < nw; ++w) { tot += w_live[w]; rays += w_rays[w]; tests += w_tests[w]; }
        block_base = tot ? atomicAdd(n_out, tot) : 0;
        if (rays) atomicAdd(ctr + bbrt::kCtrRays, rays);
        if (tests) atomicAdd(ctr + bbrt::kCtrBvh, tests);
    }
    __syncthreads();
    if (alive) {
        int off = block_base;
        for (int w = 0; w < wid; ++w) off += w_live[w];
        off += (int)__popcll(mask & ((1ull << lane) - 1ull));
        live_out[off] = r;
        if (group_bits > 0) bin_out[off] = bin;
    }
}

// Exclusive scan of the group bins' counts: start[b] = rays in bins before b.
// One block of 1024 threads, kScanChunk bins a pass: the chunk is loaded
// coalesced into LDS, thread t scans its 8 consecutive bins, the 1024 sums
// are scanned across the block, and the starts go back out coalesced; the
// running total carries into the next pass.
constexpr int kScanChunk = 8192;   // 1024 threads x 8 bins
__global__ void __launch_bounds__(1024) k_scan_bins(const int32_t* __restrict__ hist, int32_t* __restrict__ start,
                                                    int nbins) {
    __shared__ int buf[kScanChunk];
    __shared__ int part[1024];
    const int t = (int)threadIdx.x;
    int carry = 0;
    for (int base = 0; base < nbins; base += kScanChunk) {
        const int m = min(kScanChunk, nbins - base);
        for (int i = t; i < kScanChunk; i += 1024) buf[i] = i < m ? hist[base + i] : 0;
        __syncthreads();
        int v[8];
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) { v[k] = buf[t * 8 + k]; sum += v[k]; }
        part[t] = sum;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int u = t >= off ? part[t - off] : 0;
            __syncthreads();
            part[t] += u;
            __syncthreads();
        }
        int run = carry + (t ? part[t - 1] : 0);
#pragma unroll
        for (int k = 0; k < 8; ++k) { buf[t * 8 + k] = run; run += v[k]; }
        const int total = part[1023];
        __syncthreads();
        for (int i = t; i < m; i += 1024) start[base + i] = buf[i];
        carry += total;
        __syncthreads();   // the next pass overwrites buf and part
    }
}

// The live list grouped by bin: each ray to the next free place of its bin.
__global__ void k_group_rays(const int32_t* __restrict__ live, const int32_t* __restrict__ bin,
                             const int32_t* __restrict__ n_live, int32_t* cursor, int32_t* __restrict__ grouped) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= *n_live) return;
    grouped[atomicAdd(cursor + bin[i], 1)] = live[i];
}

// Persistent bounce loop as a per-lane state machine under a wave-uniform
// outer loop.  Each iteration: lanes without a ray take one from the live list
// (one wave-aggregated atomic); traversing lanes walk 4-wide nodes until they
// stand on a leaf that passes its exact test (while-while); lanes on a leaf run
// stage 1 on its entered objects until one hits (the whole wave runs each
// lane's stage 1 in turn when few lanes are in this phase); lanes whose bounce
// is complete record it, reflect, and start the next bounce at once -- a lane
// never waits for another lane's whole bounce.  Every ray ends within
// max_bounces bounces and the list is finite, so every wave leaves the loop.
enum : int { kPhTrav = 1, kPhLeaf = 2, kPhDone = 3, kPhSeed = 4 };

// A lane whose stage 1 hits parks (kPhSeed) with its triangle seed; once
// newton_batch lanes of the wave are parked, or no other lane can make
// progress, they run Newton together, so Newton's FP64 work is spread over
// many lanes instead of one or two (the results do not depend on the
// schedule).  LEARNED: at that point the wave first evaluates the network for
// the parked lanes; Newton starts from the network's seed and its result is
// kept only under the acceptance rule (converged, polished, within
// mlp::kSeedWindow of the triangle seed), else Newton runs again from the
// triangle seed and its outcome stands.
// FIRST (with LEARNED, BBRT_FLAG_SEED_FIRST): the network seed goes to Newton before stage 1, which runs only
// when that Newton neither converges inside the patch nor clearly outside it.
template <int W, bool WORK, bool LEARNED, bool COMPACT, bool F32 = false, bool FIRST = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W)))
k_persist(DevScene S, RayState R, const int32_t* __restrict__ live, const int32_t* __restrict__ n_live_p,
          int32_t* head, Outs O, int max_bounces, int record_k, int mode, uint32_t flags, unsigned long long* ctr,
          MlpDev MD, int newton_batch) {
    const MlpLds* lds = nullptr;
    if constexpr (LEARNED) {
        __shared__ MlpLds mlp_lds;
        mlp_stage_lds(MD, mlp_lds, threadIdx.x, blockDim.x);
        __syncthreads();
        lds = &mlp_lds;
    }
    Owner* slots = nullptr;   // this wave's 64 owner records for the wave-wide stage 1 (dynamic LDS)
    if constexpr (!FIRST) {
        extern __shared__ __align__(16) unsigned char bbrt_dyn_lds[];
        slots = reinterpret_cast<Owner*>(bbrt_dyn_lds) + (threadIdx.x & ~63u);
    }
    const int n_live = *n_live_p;
    const int lane = __lane_id();
    Counters c = {};
    Counters c0 = {};
    int32_t r = -1;   // -1: wants a ray, -2: list exhausted
    int scn = 0;      // the ray's sub-scene
    int b = 0, phase = 0;
    V3 o = { 0, 0, 0 }, d = { 0, 0, 0 };
    double inv[3] = { 0, 0, 0 };
    bool cull = false;
    Ray32 r32 = { 0, 0, 0, 0, 0, 0 };
    Ent4 stack[kStack4Cap];
    int sp = 0;
    int32_t cur = -1;                         // 4-wide node (>= 0), leaf (-2 - (start << 3 | count)) or nothing (-1)
    float cur_t = 0.0f;                       // its entry t (FP32, prune only)
    uint32_t negm = 0;                        // direction sign bits: the depth-first order
    double best_t = INFINITY;
    // the prune bound best_t * kPruneSlack rounded down to a float: for a float entry t,
    // (double)t > best_t * kPruneSlack exactly when t > prune_f
    float prune_f = INFINITY;
    PHit best{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
    int32_t best_obj = -1, err_obj = -1;
    int err = 0;
    int leaf_start = 0, leaf_count = 0, q = 0;
    uint32_t enter = 0;
    int32_t obj = 0;                          // the leaf object being intersected
    uint32_t sl0 = 0, sf0 = 0;                // learned mode: seed counters when the ray started
    Seed sd{ 0.0, 0.0, 0.0, false };          // its stage-1 seed, while parked for Newton

    // bounce set-up: RayTracingDriver.hpp:187 and the root of BVH.hpp:89
    auto start_bounce = [&](bool root_live) {
        cull = cull_ok(S, scn, o, d, flags);
        r32 = { (float)o.x, (float)o.y, (float)o.z, 1.0f / (float)d.x, 1.0f / (float)d.y, 1.0f / (float)d.z };
        inv[0] = 1.0 / d.x; inv[1] = 1.0 / d.y; inv[2] = 1.0 / d.z;
        best_t = INFINITY;
        prune_f = INFINITY;
        best_obj = -1; err_obj = -1; err = 0; sp = 0;
        negm = (inv[0] < 0.0 ? 1u : 0u) | (inv[1] < 0.0 ? 2u : 0u) | (inv[2] < 0.0 ? 4u : 0u);
        const int32_t ri = S.scenes[scn].root;
        double tm = -INFINITY;
        const bool ok = root_live ? true : box_static(o, inv, S.bvh[ri], tm);
        cur = ok ? ri : -1;                   // the root's 4-wide node (built for a leaf root too)
        cur_t = -INFINITY;
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
                    scn = ray_scene_of(S, r);
                    if constexpr (LEARNED) { sl0 = c.seed_learned; sf0 = c.seed_fallback; }
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

        // traversal: 4-wide nodes until a leaf to intersect, or the end (BVH.hpp:91-126).
        // Slots are tested in FP32 (a superset of the binary nodes the FP64 walk
        // enters) and visited in the binary walk's depth-first order; a leaf is
        // intersected only if its own FP64 test with the current best_t passes,
        // which is exactly the reference's condition for popping it live (the
        // ancestors' tests pass whenever the leaf's does: their boxes contain it and
        // their entry t is no larger).  So the leaves intersected, their order and
        // every record are the binary walk's.  Rays whose origin is beyond the
        // culling limit test the slots' binary nodes in FP64 instead.
        if (r >= 0 && phase == kPhTrav) {
            for (;;) {
                // inner nodes and pops until the lane reaches a leaf or its walk ends; the lanes at a leaf wait
                // and test their leaves together below (while-while: a step runs one of the two paths, not both)
                for (;;) {
                    if (cur == -1 || cur_t > prune_f) {
                        if (sp == 0) { phase = kPhDone; break; }
                        --sp;
                        cur = stack[sp].ref; cur_t = stack[sp].tmin;
                        continue;
                    }
                    if (cur <= -2) break;
                    c.bvh_tests++;
                    const bbrt::Bvh4Node* N = S.bvh4 + cur;
                    const int4 rf = *reinterpret_cast<const int4*>(N->ref);
                    const int axes = N->axes;
                    const int32_t ref[4] = { rf.x, rf.y, rf.z, rf.w };
                    float tmn[4];
                    uint32_t hm;
                    if (cull) {
                        hm = slots32(N, r32, tmn);
                    } else {
                        hm = 0;
#pragma unroll
                        for (int sl = 0; sl < 4; ++sl) {
                            tmn[sl] = 0.0f;
                            if (ref[sl] != -1) {
                                double tm = 0.0;
                                const bool hit = ref[sl] >= 0
                                    ? box_static(o, inv, S.bvh[ref[sl]], tm)
                                    : box_static6(o, inv, S.leaf_box + 6 * (size_t)((-2 - ref[sl]) >> 3),
                                                  S.leaf_box + 6 * (size_t)((-2 - ref[sl]) >> 3) + 3, tm);
                                if (hit) {
                                    hm |= 1u << sl;
                                    tmn[sl] = (float)tm;
                                }
                            }
                        }
                    }
                    hm &= (ref[0] != -1 ? 1u : 0u) | (ref[1] != -1 ? 2u : 0u) | (ref[2] != -1 ? 4u : 0u) |
                          (ref[3] != -1 ? 8u : 0u);
                    // depth-first order of the slots: the node's split axis picks the pair, each
                    // child's axis the slot within it (BVH.hpp:118-125)
                    // (slots 0-1 are the left pair, 2-3 the right one; conditional swaps order each pair, then
                    // the pairs)
                    const bool f0 = ((negm >> (axes & 3)) & 1u) != 0, fl = ((negm >> ((axes >> 2) & 3)) & 1u) != 0,
                               fr = ((negm >> ((axes >> 4) & 3)) & 1u) != 0;
                    const int32_t a0 = fl ? ref[1] : ref[0], a1 = fl ? ref[0] : ref[1];
                    const int32_t b0 = fr ? ref[3] : ref[2], b1 = fr ? ref[2] : ref[3];
                    const float ta0 = fl ? tmn[1] : tmn[0], ta1 = fl ? tmn[0] : tmn[1];
                    const float tb0 = fr ? tmn[3] : tmn[2], tb1 = fr ? tmn[2] : tmn[3];
                    const uint32_t ma = fl ? (((hm >> 1) & 1u) | ((hm & 1u) << 1)) : (hm & 3u);
                    const uint32_t mb = fr ? (((hm >> 3) & 1u) | (((hm >> 2) & 1u) << 1)) : ((hm >> 2) & 3u);
                    const int32_t R[4] = { f0 ? b0 : a0, f0 ? b1 : a1, f0 ? a0 : b0, f0 ? a1 : b1 };
                    const float T[4] = { f0 ? tb0 : ta0, f0 ? tb1 : ta1, f0 ? ta0 : tb0, f0 ? ta1 : tb1 };
                    const uint32_t pm = f0 ? (mb | (ma << 2)) : (ma | (mb << 2));
                    if (pm == 0) { cur = -1; continue; }
                    if (sp + 3 > kStack4Cap) { err = bbrt::kErrStack; phase = kPhDone; break; }
                    const int fq = __ffs(pm) - 1;
#pragma unroll
                    for (int k = 3; k >= 1; --k)
                        if (k > fq && ((pm >> k) & 1u)) { stack[sp].ref = R[k]; stack[sp].tmin = T[k]; ++sp; }
                    cur = fq == 0 ? R[0] : (fq == 1 ? R[1] : (fq == 2 ? R[2] : R[3]));
                    cur_t = fq == 0 ? T[0] : (fq == 1 ? T[1] : (fq == 2 ? T[2] : T[3]));
                }
                if (phase != kPhTrav) break;
                c.bvh_tests++;
                {
                    // a leaf: its FP64 box, object ids and culling boxes are addressed by its start,
                    // so the loads go out together
                    const int v = -2 - cur;
                    const int ls = v >> 3, lc = v & 7;
                    const double2* lb2 = reinterpret_cast<const double2*>(S.leaf_box + 6 * (size_t)ls);
                    const double2 b01 = lb2[0], b23 = lb2[1], b45 = lb2[2];
                    uint32_t en = 0;
                    if (cull) {
                        // four objects' boxes a round trip (their loads go out together)
                        const float4* b4 = reinterpret_cast<const float4*>(S.leaf_pbox + (size_t)ls * 8);
                        for (int k0 = 0; k0 < lc; k0 += 4) {
                            float4 a[4], bb[4];
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                if (k0 + u < lc) { a[u] = b4[2 * (k0 + u)]; bb[u] = b4[2 * (k0 + u) + 1]; }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                if (k0 + u < lc) {
                                    const float bx[6] = { a[u].x, a[u].y, a[u].z, a[u].w, bb[u].x, bb[u].y };
                                    if (box32(bx, r32)) en |= 1u << (k0 + u);
                                }
                            }
                        }
                    } else {
                        en = (1u << lc) - 1u;
                    }
                    const double blo[3] = { b01.x, b01.y, b23.x }, bhi[3] = { b23.y, b45.x, b45.y };
                    double tm = 0.0;
                    if (box_static6(o, inv, blo, bhi, tm) && !(tm > best_t)) {
                        if (cull) c.cull_tests += lc;
                        if (en == 0) {            // no object's culling box is entered: no stage-1 hit here
                            c.candidates += lc;
                            cur = -1;
                            continue;
                        }
                        leaf_start = ls; leaf_count = lc; enter = en;
                        q = __ffs(en) - 1;        // objects before the first entered one: counted, skipped
                        c.candidates += q;
                        phase = kPhLeaf;
                        break;
                    }
                    cur = -1;
                    continue;
                }
            }
        }

        // leaf objects: SurfacePatch::intersect and the best-hit update (BVH.hpp:101-109).  q is always an
        // entered object; stage 1 runs on the leaf's entered objects in order until one hits (the lane then
        // parks for Newton) or the leaf ends, so misses cost no extra pass of the loop.
        auto next_object = [&]() -> bool {
            ++q;
            const uint32_t rest = q < leaf_count ? (enter >> q) : 0u;
            if (rest == 0) { c.candidates += leaf_count - q; cur = -1; phase = kPhTrav; return false; }
            const int nq = q + __ffs(rest) - 1;
            c.candidates += nq - q;
            q = nq;
            return true;
        };
        bool have_h = false;
        PHit h{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
        const unsigned long long in_leaf = __ballot(r >= 0 && phase == kPhLeaf);
        if constexpr (FIRST) {
            // straight to the network and Newton with the leaf's next entered object
            if (r >= 0 && phase == kPhLeaf) {
                obj = S.leaf[leaf_start + q];
                c.candidates++;
                phase = kPhSeed;
            }
        } else if (!COMPACT && (flags & BBRT_FLAG_STAGE1_LANE)) {
            // opt-in, stored scenes: each lane walks its own pyramids, object after object
            if (r >= 0 && phase == kPhLeaf) {
                for (;;) {
                    obj = S.leaf[leaf_start + q];
                    sd = stage1_seed<0>(S, obj, o, d, r32, cull, S.scenes[scn].cull_widen, c);
                    if (sd.have) { phase = kPhSeed; break; }   // parked until the wave runs Newton
                    if (!next_object()) break;
                }
            }
        } else {
            // Every leaf lane first tests its object's nine super boxes itself, all in one round trip: an object
            // none of whose supers the ray enters has no stage-1 hit (counted as the walk counts it), and the lane
            // moves on.  Then the wave walks the pyramids below the hit supers of all its leaf lanes together,
            // breadth-first (stage1_bfs).  Objects without culling are walked by the whole wave one lane at a time
            // (stored scenes, stage1_coop; also every object under BBRT_FLAG_STAGE1_WAVE), or tile by tile with
            // the wave testing the cells (compact scenes).  The boxes, counts and results are the walk's own.
            // Leaf lanes wait (the others traverse on) until kStage1Gather of them are in a leaf or no other
            // live lane has work before Newton, so the passes hold more entries; stage 1's outcome does not
            // depend on when it runs.
            const unsigned long long parked = __ballot(r >= 0 && phase == kPhSeed);
            const bool go = in_leaf != 0 &&
                            ((int)__popcll(in_leaf) >= kStage1Gather || (in_leaf | parked) == __ballot(r >= 0));
            if (go) {
                uint32_t sup = 0;
                auto super_level = [&]() {   // from the lane's current entered object: the next one the walk must see
                    for (;;) {
                        obj = S.leaf[leaf_start + q];
                        if (!(cull && S.meta[obj].cull)) { sup = 0; return; }   // no culling: the walk does it
                        sup = group_mask<9, bbrt::kG9>(S.box_super + (size_t)obj * bbrt::kSuperFloats, r32);
                        c.cull_tests += 9;
                        if (sup) return;
                        c.candidates++;
                        if (!next_object()) return;
                    }
                };
                bool want_super = r >= 0 && phase == kPhLeaf;   // one call site: the lambda is inlined once
                const bool serial_all = !COMPACT && (flags & BBRT_FLAG_STAGE1_WAVE) != 0;
                for (;;) {
                    if (want_super) { super_level(); want_super = false; }
                    const bool mine = r >= 0 && phase == kPhLeaf;
                    const unsigned long long need = __ballot(mine);
                    if (!need) break;
                    const unsigned long long serial = serial_all ? need : __ballot(mine && sup == 0);
                    Stage1 b1{ 0.0, 0, false };
                    uint32_t n_cull = 0, n_mt = 0;
                    bool done = mine;   // the lanes whose object's stage 1 ends in this pass
                    if (!serial) {
                        stage1_bfs<COMPACT ? 1 : 0>(S, mine, obj, o, d, r32, S.scenes[scn].cull_widen, sup, slots, b1,
                                                    n_cull, n_mt);
                    } else if constexpr (COMPACT) {
                        // objects without culling: each lane walks every tile, the wave tests the cells
                        bool walking = mine && sup == 0;
                        done = walking;
                        CompactWalk w{ 0u, 0u, 0u, 0, 0, 0, 0, 0, false };
                        if (walking) compact_walk_init(S, obj, r32, false, w, c);
                        while (__ballot(walking)) {
                            int ti = 0, tj = 0;
                            uint32_t mc = 0;
                            bool tile = false;
                            if (walking) { tile = compact_walk_next(S, obj, r32, w, ti, tj, mc, c); walking = tile; }
                            compact_cells_wave(S, tile, obj, ti, tj, mc, o, d, r32, S.scenes[scn].cull_widen, false,
                                               b1, c, slots);
                        }
                    } else {
                        const int ld = __ffsll((long long)serial) - 1;
                        const int32_t kk = __shfl(obj, ld);
                        const uint32_t lsup = (uint32_t)__shfl((int)sup, ld);
                        const V3 lo_ = { __shfl(o.x, ld), __shfl(o.y, ld), __shfl(o.z, ld) };
                        const V3 ld_ = { __shfl(d.x, ld), __shfl(d.y, ld), __shfl(d.z, ld) };
                        const Ray32 lr = { __shfl(r32.ox, ld), __shfl(r32.oy, ld), __shfl(r32.oz, ld),
                                           __shfl(r32.ix, ld), __shfl(r32.iy, ld), __shfl(r32.iz, ld) };
                        const int lscn = __shfl(scn, ld);
                        const bool lcull = __shfl(cull ? 1 : 0, ld) != 0;
                        b1 = stage1_coop<0>(S, kk, lo_, ld_, lr, S.scenes[lscn].cull_widen, lcull && S.meta[kk].cull,
                                            lsup, n_cull, n_mt);
                        done = lane == ld;
                    }
                    if (done) {
                        c.candidates++;
                        c.cull_tests += n_cull;
                        c.mt_tests += n_mt;
                        if (b1.have) { sd = seed_of(b1); phase = kPhSeed; }   // parked until the wave runs Newton
                        else if (next_object()) want_super = true;
                    }
                }
            }
        }
        // Newton for the parked lanes, all together: once newton_batch lanes wait, or
        // when no other lane can make progress
        const unsigned long long pend = __ballot(r >= 0 && phase == kPhSeed);
        const unsigned long long busy = __ballot((r >= 0 && phase != kPhSeed) || r == -1);
        if (pend && ((int)__popcll(pend) >= newton_batch || busy == 0)) {
            const bool need = r >= 0 && phase == kPhSeed;
            if constexpr (LEARNED && !FIRST) {
                // Newton from the network's seed; lanes whose root fails the acceptance rule go round once
                // more from the triangle seed.  One call site: the solver is inlined once.
                double ul, vl, tl;
                mlp_seed_wave(MD, *lds, need, obj, o, d, ul, vl, tl);
                if (lane == 0) c.mlp_evals++;
                if (need) c.mlp_lanes++;
                bool again = need;
                for (int attempt = 0; __ballot(again) != 0; ++attempt) {
                    if (again) {
                        h = newton_any<F32>(S, obj, o, d, ul, vl, tl, mode, c);
                        again = false;
                        if (attempt == 0) {
                            const bool ok = h.status == 1 && h.polish_ok && fabs(h.u - sd.u) <= mlp::kSeedWindow &&
                                            fabs(h.v - sd.v) <= mlp::kSeedWindow;
                            if (ok) {
                                c.seed_learned++;
                            } else {
                                c.seed_fallback++;
                                ul = sd.u; vl = sd.v; tl = sd.t;
                                again = true;
                            }
                        }
                    }
                }
            } else if constexpr (LEARNED) {   // FIRST: the network seed before stage 1
                double ul, vl, tl;
                mlp_seed_wave(MD, *lds, need, obj, o, d, ul, vl, tl);
                if (lane == 0) c.mlp_evals++;
                if (need) {
                    c.mlp_lanes++;
                    h = newton_any<F32, FIRST>(S, obj, o, d, ul, vl, tl, mode, c);
                    constexpr double lim = 1.0 + BBRT_SEED_FIRST_MISS_MARGIN;
                    // a root behind the origin or at it (the point the ray was nudged off) is no stage-1
                    // hit: Moller-Trumbore rejects t < 1e-9 (MollerTrumbore.hpp:54-55)
                    if (h.status == 1 && h.polish_ok && h.t >= kMtEps) {
                        c.seed_learned++;
                    } else if (h.status == -bbrt::kErrDomain && (fabs(h.u) > lim || fabs(h.v) > lim)) {
                        c.seed_miss++;
                        h = PHit{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
                    } else {
                        // stage 1 and the triangle seed decide (SurfacePatch.hpp:253-303)
                        c.seed_fallback++;
                        c.candidates--;   // stage1_seed counts the candidate again
                        sd = stage1_seed<COMPACT ? 1 : 0>(S, obj, o, d, r32, cull, S.scenes[scn].cull_widen, c);
                        h = sd.have ? newton_any<F32>(S, obj, o, d, sd.u, sd.v, sd.t, mode, c)
                                    : PHit{ 0, 0.0, { 0, 0, 0 }, { 0, 0, 0 }, 0 };
                    }
                }
            } else {
                if (need) h = newton_any<F32>(S, obj, o, d, sd.u, sd.v, sd.t, mode, c);
            }
            if (need) { phase = kPhLeaf; have_h = true; }
        }
        if (have_h) {
            if (h.status < 0) { err = -h.status; err_obj = obj; phase = kPhDone; }
            else if (h.status == 1 && h.t < best_t) {
                best_t = h.t; best = h; best_obj = obj;
                prune_f = round_down_f(best_t * kPruneSlack);
            }
            if (phase == kPhLeaf) next_object();   // the next entered object, or back to the walk
        }

        if (r >= 0 && phase == kPhDone) {
            const bool alive = finish_bounce(S, O, r, scn, o, d, b, max_bounces, record_k, flags, err, err_obj, best,
                                             best_obj, c);
            if constexpr (WORK) add_ray_work(O, r, c, c0);
            if constexpr (LEARNED)
                if (!alive && O.seed_stats) {
                    O.seed_stats[2 * (size_t)r] = c.seed_learned - sl0;
                    O.seed_stats[2 * (size_t)r + 1] = c.seed_fallback - sf0;
                }
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

// Culling boxes in leaf order: leaf_pbox[p] = box_patch[leaf[p]] (two float4 per box).
__global__ void k_gather_leaf_pbox(const int32_t* __restrict__ leaf, const float* __restrict__ box_patch, int64_t n,
                                   float* __restrict__ leaf_pbox) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 2 * n) return;
    const int64_t p = g >> 1, h = g & 1;
    reinterpret_cast<float4*>(leaf_pbox)[2 * p + h] = reinterpret_cast<const float4*>(box_patch)[2 * (int64_t)leaf[p] + h];
}

__global__ void k_zero_i32(int32_t* p, int n) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = 0;
}

// ---------------------------------------------------------------------------
// BounceMapDriver::run on the device (BounceMapDriver.hpp:153-238)
// ---------------------------------------------------------------------------
// One axis of the first-form interpolation operator at x (PatchInterpolation.hpp:54-159, as the host's
// interpolate_2d forms it): r[k] = w_k / (x - node_k) and scale = prod_k (x - node_k), or a one-hot row
// with scale 1 within 1e-12 of a node.
__device__ __forceinline__ double ff_row(const double* __restrict__ nodes, const double* __restrict__ w, int n,
                                         double x, double* r) {
    double prod = 1.0;
    int hot = -1;
    for (int k = 0; k < n; ++k) {
        const double diff = x - nodes[k];
        if (fabs(diff) < kOnNodeTol) { hot = k; break; }
        r[k] = w[k] / diff;
        prod *= diff;
    }
    if (hot >= 0) {
        for (int k = 0; k < n; ++k) r[k] = 0.0;
        r[hot] = 1.0;
        return 1.0;
    }
    return prod;
}

// SurfacePatch::sample_grid (SurfacePatch.hpp:199-242) at grid point (i, j) of an n x n grid: the first-form
// value of a field, with the two products summed over their inner index upwards from 0.0.
__device__ __forceinline__ double ff_value(const double* rx, double sx, const double* ry, double sy,
                                           const double* __restrict__ f, int im, int jm) {
    double acc = 0.0;
    for (int m = 0; m < jm; ++m) {
        double mid = 0.0;
        for (int k = 0; k < im; ++k) mid += rx[k] * f[(size_t)k * jm + m];
        acc += mid * ry[m];
    }
    return acc * (sx * sy);
}

// One thread per (sampled patch, grid point): the grid point, its twice-normalised normal (sample_grid, then
// run()), the tangent frame (build_tangent_frame, BounceMapDriver.hpp:499-519), the origin nudged off the
// surface by the patch's self-intersection epsilon, and the theta_n x phi_n cone of launch directions
// (BounceMapDriver.hpp:202-220), rays point-major, theta then phi.  cos(phi) and sin(phi) come from the host
// (its libm), so the rays are the host computation's bit for bit.
__global__ void k_bmap_launch(DevScene S, const int64_t* __restrict__ pids, const double* __restrict__ eps,
                              int64_t n_sample, int gn, int theta_n, int phi_n, double cos_theta_max,
                              const double* __restrict__ cphi, const double* __restrict__ sphi, double* grid,
                              double* o_out, double* d_out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = (int64_t)gn * gn;
    if (g >= n_sample * nn) return;
    const int64_t ps = g / nn;
    const int k = (int)(g - ps * nn);
    const int i = k / gn, j = k - (k / gn) * gn;
    const PatchMeta M = S.meta[pids[ps]];
    const int im = M.imax, jm = M.jmax;
    const double* un = S.nodes + M.node_off;
    const double* uw = un + im;
    const double* vn = uw + im;
    const double* vw = vn + jm;
    const double* F = S.fields + M.field_off;
    const int nf = im * jm;
    const double ui = -1.0 + 2.0 * static_cast<double>(i) / (gn - 1);
    const double vj = -1.0 + 2.0 * static_cast<double>(j) / (gn - 1);
    double rx[bbrt::kMaxNodes1D], ry[bbrt::kMaxNodes1D];
    const double sx = ff_row(un, uw, im, ui, rx);
    const double sy = ff_row(vn, vw, jm, vj, ry);
    // fields: x y z (0-2), dX/du (3-5), normal (9-11)
    const V3 p = { ff_value(rx, sx, ry, sy, F, im, jm), ff_value(rx, sx, ry, sy, F + nf, im, jm),
                   ff_value(rx, sx, ry, sy, F + 2 * nf, im, jm) };
    const V3 du = { ff_value(rx, sx, ry, sy, F + 3 * nf, im, jm), ff_value(rx, sx, ry, sy, F + 4 * nf, im, jm),
                    ff_value(rx, sx, ry, sy, F + 5 * nf, im, jm) };
    const V3 ng = unit3({ ff_value(rx, sx, ry, sy, F + 9 * nf, im, jm), ff_value(rx, sx, ry, sy, F + 10 * nf, im, jm),
                          ff_value(rx, sx, ry, sy, F + 11 * nf, im, jm) });
    if (grid) {
        double* q = grid + 9 * g;
        q[0] = p.x; q[1] = p.y; q[2] = p.z; q[3] = ng.x; q[4] = ng.y; q[5] = ng.z; q[6] = du.x; q[7] = du.y; q[8] = du.z;
    }
    const V3 n = unit3(ng);
    // build_tangent_frame
    const double du_len2 = dot3(du, du);
    const double du_dot_n = dot3(du, n);
    const V3 du_perp = { du.x - du_dot_n * n.x, du.y - du_dot_n * n.y, du.z - du_dot_n * n.z };
    const double du_perp_len2 = dot3(du_perp, du_perp);
    V3 t1;
    if (du_perp_len2 > 1e-6 * 1e-6 * du_len2) {
        t1 = unit3(du_perp);
    } else {
        const V3 a = (fabs(n.x) <= fabs(n.y) && fabs(n.x) <= fabs(n.z)) ? V3{ 1.0, 0.0, 0.0 }
                     : (fabs(n.y) <= fabs(n.z) ? V3{ 0.0, 1.0, 0.0 } : V3{ 0.0, 0.0, 1.0 });
        t1 = unit3(cross3(n, a));
    }
    const V3 t2 = cross3(n, t1);
    const double e = eps[ps];
    const V3 org = { p.x + n.x * e, p.y + n.y * e, p.z + n.z * e };
    if (!o_out) return;
    int64_t r = g * theta_n * phi_n;
    for (int ti = 0; ti < theta_n; ++ti) {
        const double mu = cos_theta_max + (static_cast<double>(ti) + 0.5) / static_cast<double>(theta_n) *
                                              (1.0 - cos_theta_max);
        const double sq = 1.0 - mu * mu;
        const double st = sqrt(0.0 < sq ? sq : 0.0);
        for (int pj = 0; pj < phi_n; ++pj, ++r) {
            const double cp = cphi[pj], sp = sphi[pj];
            o_out[3 * r] = org.x; o_out[3 * r + 1] = org.y; o_out[3 * r + 2] = org.z;
            d_out[3 * r] = t1.x * st * cp + t2.x * st * sp + n.x * mu;
            d_out[3 * r + 1] = t1.y * st * cp + t2.y * st * sp + n.y * mu;
            d_out[3 * r + 2] = t1.z * st * cp + t2.z * st * sp + n.z * mu;
        }
    }
}

// Per grid point over its directions (BounceMapDriver.hpp:214-236): Error directions are counted and
// skipped; the largest bounce count of the others, and the first direction that reaches it (strictly larger
// than the best so far).  best_od: that direction's launch ray.
__global__ void k_bmap_reduce(const int8_t* __restrict__ status, const int32_t* __restrict__ bounces,
                              const double* __restrict__ o, const double* __restrict__ d, int64_t n_points,
                              int per, int32_t* max_b, int32_t* err_c, int32_t* best_dir, double* best_od) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_points) return;
    int best = 0, n_err = 0, arg = -1, arg_b = 0;
    for (int q = 0; q < per; ++q) {
        const int64_t r = g * per + q;
        if (status[r] == bbrt::kError) { ++n_err; continue; }
        const int b = bounces[r];
        if (arg < 0 || b > arg_b) { arg = q; arg_b = b; }
        best = best < b ? b : best;
    }
    if (max_b) max_b[g] = best;
    if (err_c) err_c[g] = n_err;
    if (best_dir) best_dir[g] = arg;
    if (best_od) {
        double* q = best_od + 6 * g;
        const int64_t r = g * per + (arg < 0 ? 0 : arg);
        for (int c = 0; c < 3; ++c) { q[c] = arg < 0 ? 0.0 : o[3 * r + c]; q[3 + c] = arg < 0 ? 0.0 : d[3 * r + c]; }
    }
}

// Device buffers of one call, freed when it returns.
struct CallBuffers {
    std::vector<void*> p;
    ~CallBuffers() { for (void* q : p) (void)hipFree(q); }
    template <class T>
    T* get(size_t count) {
        void* q = nullptr;
        check(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
        p.push_back(q);
        return static_cast<T*>(q);
    }
};

}  // namespace

// Scratch of the no-allocation entry point, bbrt_trace_device_noalloc (file scope;
// each GPU holds its own copy at its own address until the process ends).
static __device__ __align__(256) unsigned char scratch[BBRT_NOALLOC_SCRATCH_BYTES];

namespace {

// ---------------------------------------------------------------------------
// Device build of compact scenes (bit-identical to scene_build.cpp)
// ---------------------------------------------------------------------------
// Pass 1, one workgroup per patch: max sagitta, finiteness of every coarse
// value, longest coarse-cell diagonal (prepare_patch's statistics).
__global__ void __launch_bounds__(256) k_patch_stats(DevScene S, double* __restrict__ max_sag, int32_t* __restrict__ finite,
                                                     double* __restrict__ hmax) {
    __shared__ double red_ms[256], red_h[256];
    __shared__ int red_f[256];
    const int64_t k = blockIdx.x;
    double ms = 0.0, h = 0.0;
    bool fin = true;
    for (int c = threadIdx.x; c < bbrt::kCells * bbrt::kCells; c += blockDim.x) {
        const int i = c / bbrt::kCells, j = c % bbrt::kCells;
        const CellData cd = compact_cell_any(S, k, i, j);
        ms = (ms < cd.s) ? cd.s : ms;
        const V3* P[4] = { &cd.p00, &cd.p10, &cd.p01, &cd.p11 };
        bool f = isfinite(cd.s) && isfinite(cd.n.x) && isfinite(cd.n.y) && isfinite(cd.n.z);
        for (int q = 0; q < 4; ++q) f = f && isfinite(P[q]->x) && isfinite(P[q]->y) && isfinite(P[q]->z);
        fin = fin && f;
        const double ax = cd.p11.x - cd.p00.x, ay = cd.p11.y - cd.p00.y, az = cd.p11.z - cd.p00.z;
        const double ex = cd.p10.x - cd.p01.x, ey = cd.p10.y - cd.p01.y, ez = cd.p10.z - cd.p01.z;
        double a = 0.0, e = 0.0;
        a += ax * ax; a += ay * ay; a += az * az;
        e += ex * ex; e += ey * ey; e += ez * ez;
        const double hc = sqrt((a < e) ? e : a);
        h = (h < hc) ? hc : h;
    }
    red_ms[threadIdx.x] = ms; red_h[threadIdx.x] = h; red_f[threadIdx.x] = fin ? 1 : 0;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const double a = red_ms[threadIdx.x + off], b = red_h[threadIdx.x + off];
            red_ms[threadIdx.x] = (red_ms[threadIdx.x] < a) ? a : red_ms[threadIdx.x];
            red_h[threadIdx.x] = (red_h[threadIdx.x] < b) ? b : red_h[threadIdx.x];
            red_f[threadIdx.x] &= red_f[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        max_sag[k] = red_ms[0];
        finite[k] = red_f[0];
        hmax[k] = red_f[0] ? red_h[0] : 0.0;
    }
}

// Pass 2, one workgroup per patch of one sub-scene: the FP32 pyramid in the
// grouped layout (build_pyramid without the cell level, which compact scenes
// rebuild on the fly).
__global__ void __launch_bounds__(256) k_patch_pyramid(DevScene S, int64_t p0, double widen, float* __restrict__ box_tile,
                                                       float* __restrict__ box_block, float* __restrict__ box_super,
                                                       float* __restrict__ box_patch, uint8_t* __restrict__ qbox) {
    __shared__ float tb[bbrt::kTileN * bbrt::kTileN][6];
    __shared__ float bb[bbrt::kBlockN * bbrt::kBlockN][6];
    __shared__ float sb[bbrt::kSuperN * bbrt::kSuperN][6];
    const int64_t k = p0 + blockIdx.x;
    const int tid = threadIdx.x;
    const bool cull = S.meta[k].cull != 0;
    float* gt = box_tile + (size_t)k * bbrt::kTileFloats;
    float* gb = box_block + (size_t)k * bbrt::kBlockFloats;
    float* gs = box_super + (size_t)k * bbrt::kSuperFloats;
    float* gp = box_patch + (size_t)k * 8;
    auto merge = [](float* dst, const float* src) {
        for (int a = 0; a < 3; ++a) { dst[a] = (src[a] < dst[a]) ? src[a] : dst[a]; dst[3 + a] = (dst[3 + a] < src[3 + a]) ? src[3 + a] : dst[3 + a]; }
    };
    auto empty = [](float* b) { b[0] = b[1] = b[2] = INFINITY; b[3] = b[4] = b[5] = -INFINITY; };
    if (cull) {
        for (int t = tid; t < bbrt::kTileN * bbrt::kTileN; t += blockDim.x) {
            const int ti = t / bbrt::kTileN, tj = t % bbrt::kTileN;
            float b[6];
            float cb[9][6];
            empty(b);
            for (int cq = 0; cq < 9; ++cq) {
                const CellData cd = compact_cell_any(S, k, 3 * ti + cq / 3, 3 * tj + cq % 3);
                compact_cell_box_f(cd, widen, cb[cq]);
                merge(b, cb[cq]);
            }
            for (int a = 0; a < 6; ++a) tb[t][a] = b[a];
            uint8_t* q = qbox + (size_t)k * bbrt::kQCellPatchBytes + (size_t)t * bbrt::kQCellBytes;
            for (int cq = 0; cq < 9; ++cq) bbrt::quantize_cell_box(b, cb[cq], q + 6 * cq);
            q[54] = 0; q[55] = 0;
        }
        __syncthreads();
        for (int q = tid; q < bbrt::kBlockN * bbrt::kBlockN; q += blockDim.x) {
            const int bi = q / bbrt::kBlockN, bj = q % bbrt::kBlockN;
            float b[6];
            empty(b);
            for (int ti = 3 * bi; ti < 3 * bi + 3; ++ti)
                for (int tj = 3 * bj; tj < 3 * bj + 3; ++tj) merge(b, tb[ti * bbrt::kTileN + tj]);
            for (int a = 0; a < 6; ++a) bb[q][a] = b[a];
        }
        __syncthreads();
        if (tid < bbrt::kSuperN * bbrt::kSuperN) {
            const int si = tid / bbrt::kSuperN, sj = tid % bbrt::kSuperN;
            float b[6];
            empty(b);
            for (int bi = bbrt::kBlocksPerSuper * si; bi < min(bbrt::kBlockN, bbrt::kBlocksPerSuper * (si + 1)); ++bi)
                for (int bj = bbrt::kBlocksPerSuper * sj; bj < min(bbrt::kBlockN, bbrt::kBlocksPerSuper * (sj + 1)); ++bj)
                    merge(b, bb[bi * bbrt::kBlockN + bj]);
            for (int a = 0; a < 6; ++a) sb[tid][a] = b[a];
        }
        __syncthreads();
    }
    // tile groups: by block, slot 3*ti' + tj'; two zero pad floats per group
    for (int q = tid; q < (int)bbrt::kTileFloats; q += blockDim.x) {
        const int g = q / bbrt::kG9, w = q % bbrt::kG9;
        float v = 0.0f;
        if (w < 54) {
            const int slot = w / 6, a = w % 6;
            const int ti = 3 * (g / bbrt::kBlockN) + slot / 3, tj = 3 * (g % bbrt::kBlockN) + slot % 3;
            v = cull ? tb[ti * bbrt::kTileN + tj][a] : (a < 3 ? INFINITY : -INFINITY);
        }
        gt[q] = v;
    }
    // block groups: by super, slot 4*bi' + bj'; empty where the grid ends
    for (int q = tid; q < (int)bbrt::kBlockFloats; q += blockDim.x) {
        const int g = q / bbrt::kG16, w = q % bbrt::kG16;
        const int slot = w / 6, a = w % 6;
        const int bi = bbrt::kBlocksPerSuper * (g / bbrt::kSuperN) + slot / bbrt::kBlocksPerSuper;
        const int bj = bbrt::kBlocksPerSuper * (g % bbrt::kSuperN) + slot % bbrt::kBlocksPerSuper;
        gb[q] = (cull && bi < bbrt::kBlockN && bj < bbrt::kBlockN) ? bb[bi * bbrt::kBlockN + bj][a]
                                                                 : (a < 3 ? INFINITY : -INFINITY);
    }
    for (int q = tid; q < (int)bbrt::kSuperFloats; q += blockDim.x) {
        float v = 0.0f;
        if (q < 54) v = cull ? sb[q / 6][q % 6] : ((q % 6) < 3 ? INFINITY : -INFINITY);
        gs[q] = v;
    }
    if (tid == 0) {
        float b[6];
        if (cull) {
            empty(b);
            for (int q = 0; q < bbrt::kSuperN * bbrt::kSuperN; ++q) merge(b, sb[q]);
        } else {   // full box: always entered (such a patch is tested exhaustively)
            b[0] = b[1] = b[2] = -INFINITY; b[3] = b[4] = b[5] = INFINITY;
        }
        for (int a = 0; a < 6; ++a) gp[a] = b[a];
        gp[6] = 0.0f; gp[7] = 0.0f;
    }
}

// The counters of an asynchronous trace, into the caller's device array
// (zeros when ctr is NULL: an empty trace).
__global__ void k_counters_out(const unsigned long long* __restrict__ ctr, int64_t* __restrict__ out, int64_t launches) {
    const int q = (int)threadIdx.x;
    if (q < BBRT_NCOUNTERS) out[q] = q == bbrt::kCtrLaunches ? launches : (ctr ? (int64_t)ctr[q] : 0);
}

// Per-ray sub-scene ids from the sub-scene ray ranges (binary search).
__global__ void k_ray_scene(const int64_t* __restrict__ begin, int n_scenes, int64_t n, int32_t* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int lo = 0, hi = n_scenes - 1;   // last sub-scene whose range starts at or before g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= g) lo = mid; else hi = mid - 1;
    }
    out[g] = lo;
}

// The no-allocation entry point's scratch (bbrt_trace_device_noalloc): ray state,
// live lists, bins, counts and counters of one trace, in a file-scope device array
// that each GPU holds until the process ends.  Host code never passes `scratch`
// itself to a kernel (in host code the name is a stand-in, not a device address):
// noalloc_scratch() takes the array's address on the current device with
// hipGetSymbolAddress once per device and keeps it.
constexpr size_t kNoAllocScratchBytes = BBRT_NOALLOC_SCRATCH_BYTES;

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
char* g_noalloc_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current device.
char* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("bbrt_trace_device_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("bbrt_trace_device_noalloc: no device address for the scratch array");
        g_noalloc_ptr[device] = static_cast<char*>(p);
    }
    return g_noalloc_ptr[device];
}

// async: enqueue on st and return without waiting (bbrt_trace_device_async):
// the stream's own scratch, counters (if d_counters) written on the device.
// fixed_scratch (with async): a caller-owned scratch of kNoAllocScratchBytes
// instead of the handle's; nothing is allocated.
void trace_impl(bbrt_scene s, const double* d_o, const double* d_d, int64_t n, const bbrt_params* p,
                const bbrt_outputs* out, int64_t* counters, hipStream_t st, double* kernel_ms, bool pad_records,
                const int32_t* d_ray_scene, bool async = false, int64_t* d_counters = nullptr,
                char* fixed_scratch = nullptr) {
    if (!p) throw std::runtime_error("bbrt_trace: params is NULL");
    if (async && (p->flags & (BBRT_FLAG_WAVEFRONT | BBRT_FLAG_GRAPH | BBRT_FLAG_TIME_KERNELS)))
        throw std::runtime_error("bbrt_trace_device_async: BBRT_FLAG_WAVEFRONT, BBRT_FLAG_GRAPH and "
                                 "BBRT_FLAG_TIME_KERNELS need the synchronous entry points");
    if (s->n_scenes > 1 && !d_ray_scene)
        throw std::runtime_error("bbrt_trace: this scene holds several sub-scenes; use bbrt_trace_batch");
    DevScene dev = s->dev;
    dev.ray_scene = d_ray_scene;
    if (p->max_bounces < 0 || p->record_k < 0) throw std::runtime_error("bbrt_trace: negative max_bounces or record_k");
    if (p->newton_mode != BBRT_NEWTON_REF && p->newton_mode != BBRT_NEWTON_CONVERGED &&
        p->newton_mode != BBRT_NEWTON_CONVERGED_FP32)
        throw std::runtime_error("bbrt_trace: unknown newton_mode");
    if (p->seed_mode != BBRT_SEED_TRIANGLE && p->seed_mode != BBRT_SEED_LEARNED)
        throw std::runtime_error("bbrt_trace: unknown seed_mode");
    const bool learned = p->seed_mode == BBRT_SEED_LEARNED;
    if (learned) {
        if (!s->mlp_ready) throw std::runtime_error("bbrt_trace: the learned seed needs a network (bbrt_scene_set_mlp)");
        if (p->newton_mode == BBRT_NEWTON_REF)
            throw std::runtime_error("bbrt_trace: the learned seed needs BBRT_NEWTON_CONVERGED (or its FP32 start)");
        if (p->flags & BBRT_FLAG_WAVEFRONT)
            throw std::runtime_error("bbrt_trace: the learned seed runs in the persistent schedule only");
        if (out && out->ray_work) throw std::runtime_error("bbrt_trace: ray_work is not recorded with the learned seed");
    }
    if (n < 0 || n > (int64_t)INT32_MAX - 1024) throw std::runtime_error("bbrt_trace: ray count out of range");
    const int block = p->block > 0 ? p->block : default_block(learned, s->compact != 0);
    if (block % 64 != 0 || block > 256) throw std::runtime_error("bbrt_trace: block must be a multiple of 64 <= 256");
    if (!st) st = s->stream;
    check(hipSetDevice(s->device), "hipSetDevice");
    int64_t local_ctr[BBRT_NCOUNTERS] = { 0 };
    if (n == 0) {
        if (async) {
            if (d_counters) hipLaunchKernelGGL(k_counters_out, dim3(1), dim3(64), 0, st, nullptr, d_counters, (int64_t)0);
            check(hipGetLastError(), "k_counters_out");
            return;
        }
        if (counters) std::memcpy(counters, local_ctr, sizeof(local_ctr));
        if (kernel_ms) { kernel_ms[0] = 0; kernel_ms[1] = 0; kernel_ms[2] = 0; }
        return;
    }
    // scratch: state o,d (48 B/ray), two live lists, counts, counters
    const size_t off_o = 0;
    const size_t off_d = align256(off_o + (size_t)n * 24);
    const size_t off_la = align256(off_d + (size_t)n * 24);
    const size_t off_lb = align256(off_la + (size_t)n * 4);
    // BBRT_FLAG_GROUP_RAYS (persistent schedule): bins per list position, bin counts, bin cursors
    int group_bits = 0;
    int64_t nbins = 0;
    if ((p->flags & BBRT_FLAG_GROUP_RAYS) && !(p->flags & BBRT_FLAG_WAVEFRONT) && p->max_bounces > 0) {
        group_bits = s->n_scenes > 1 ? kGroupBitsBatch : kGroupBitsOne;
        nbins = (int64_t)s->n_scenes << (3 * group_bits);
        if (nbins > ((int64_t)1 << 22)) { group_bits = 0; nbins = 0; }   // too many sub-scenes: input order
    }
    const size_t off_bin = align256(off_lb + (size_t)n * 4);
    const size_t off_hist = align256(off_bin + (group_bits ? (size_t)n * 4 : 0));
    const size_t off_cur = align256(off_hist + (size_t)nbins * 4);
    const size_t off_cnt = align256(off_cur + (size_t)nbins * 4);
    const size_t off_ctr = align256(off_cnt + 256);
    const size_t need = align256(off_ctr + BBRT_NCOUNTERS * 8);
    if (fixed_scratch && need > kNoAllocScratchBytes)
        throw std::runtime_error("bbrt_trace_device_noalloc: " + std::to_string(n) + " rays need " +
                                 std::to_string(need) + " bytes of scratch, more than the " +
                                 std::to_string(kNoAllocScratchBytes) + " of the fixed scratch array");
    char* base = fixed_scratch ? fixed_scratch
                               : static_cast<char*>(async ? async_scratch_for(s, st, need) : scratch_for(s, need));
    RayState R{ reinterpret_cast<double*>(base + off_o), reinterpret_cast<double*>(base + off_d) };
    int32_t* live[2] = { reinterpret_cast<int32_t*>(base + off_la), reinterpret_cast<int32_t*>(base + off_lb) };
    int32_t* cnt = reinterpret_cast<int32_t*>(base + off_cnt);   // [0], [1]: list sizes, [2]: queue head
    int32_t* bins = reinterpret_cast<int32_t*>(base + off_bin);
    int32_t* hist = reinterpret_cast<int32_t*>(base + off_hist);
    int32_t* cursor = reinterpret_cast<int32_t*>(base + off_cur);
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(base + off_ctr);
    Outs O{ out ? out->status : nullptr, out ? out->err_kind : nullptr, out ? out->bounces : nullptr,
            out ? out->err_patch : nullptr, out ? out->hit_patch : nullptr, out ? out->hit_xyz : nullptr,
            out ? out->newton_iters : nullptr, out ? out->final_od : nullptr,
            out ? reinterpret_cast<unsigned long long*>(out->hit_count) : nullptr, out ? out->ray_work : nullptr,
            out ? out->seed_stats : nullptr };
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
    const bool graph = (p->flags & BBRT_FLAG_GRAPH) != 0;
    if (graph && wavefront) throw std::runtime_error("bbrt_trace: BBRT_FLAG_GRAPH needs the persistent schedule");
    // the persistent kernel and its grid (host-side choices, fixed in a captured graph)
    decltype(&k_persist<2, false, false, false>) kern = nullptr;
    const bool f32 = p->newton_mode == BBRT_NEWTON_CONVERGED_FP32;
    if (f32 && (p->flags & BBRT_FLAG_WAVEFRONT))
        throw std::runtime_error("bbrt_trace: BBRT_NEWTON_CONVERGED_FP32 runs in the persistent schedule only");
    if (f32 && out && out->ray_work)
        throw std::runtime_error("bbrt_trace: ray_work is not recorded with BBRT_NEWTON_CONVERGED_FP32");
    const bool first = (p->flags & BBRT_FLAG_SEED_FIRST) != 0;
    if (first && (!learned || f32 || (p->flags & BBRT_FLAG_WAVEFRONT)))
        throw std::runtime_error("bbrt_trace: BBRT_FLAG_SEED_FIRST needs the learned seed, BBRT_NEWTON_CONVERGED and "
                                 "the persistent schedule");
    unsigned grid = 1;
    size_t lds_bytes = 0;   // dynamic LDS of the persistent kernel
    const MlpDev MD{ s->mlp_frame, s->mlp_h1pre, static_cast<const u32x2_t*>(s->mlp_wfrag), s->mlp_params };
    const int nbatch = p->newton_batch > 0 ? std::min(p->newton_batch, 64) : default_newton_batch(s->compact != 0);
    if (!wavefront && p->max_bounces > 0) {
        const int wpe = p->waves_per_eu > 0 ? p->waves_per_eu : default_waves_per_eu(s->compact);
        const bool work = O.ray_work != nullptr;
        if (wpe != 2 && wpe != 3 && wpe != 4) throw std::runtime_error("bbrt_trace: waves_per_eu must be 0, 2, 3 or 4");
        if (s->compact && work) throw std::runtime_error("bbrt_trace: ray_work is not recorded on a compact scene");
        if (first) {
            if (s->compact)
                kern = wpe == 2 ? &k_persist<2, false, true, true, false, true>
                                : (wpe == 3 ? &k_persist<3, false, true, true, false, true>
                                            : &k_persist<4, false, true, true, false, true>);
            else
                kern = wpe == 2 ? &k_persist<2, false, true, false, false, true>
                                : (wpe == 3 ? &k_persist<3, false, true, false, false, true>
                                            : &k_persist<4, false, true, false, false, true>);
        } else if (f32) {
            if (s->compact && learned)
                kern = wpe == 2 ? &k_persist<2, false, true, true, true>
                                : (wpe == 3 ? &k_persist<3, false, true, true, true> : &k_persist<4, false, true, true, true>);
            else if (s->compact)
                kern = wpe == 2 ? &k_persist<2, false, false, true, true>
                                : (wpe == 3 ? &k_persist<3, false, false, true, true> : &k_persist<4, false, false, true, true>);
            else if (learned)
                kern = wpe == 2 ? &k_persist<2, false, true, false, true>
                                : (wpe == 3 ? &k_persist<3, false, true, false, true> : &k_persist<4, false, true, false, true>);
            else
                kern = wpe == 2 ? &k_persist<2, false, false, false, true>
                                : (wpe == 3 ? &k_persist<3, false, false, false, true> : &k_persist<4, false, false, false, true>);
        } else if (s->compact) {
            if (learned) kern = wpe == 2 ? &k_persist<2, false, true, true> : (wpe == 3 ? &k_persist<3, false, true, true>
                                                                                        : &k_persist<4, false, true, true>);
            else kern = wpe == 2 ? &k_persist<2, false, false, true> : (wpe == 3 ? &k_persist<3, false, false, true>
                                                                                  : &k_persist<4, false, false, true>);
        } else if (learned) {
            kern = wpe == 2 ? &k_persist<2, false, true, false> : (wpe == 3 ? &k_persist<3, false, true, false>
                                                                             : &k_persist<4, false, true, false>);
        } else if (work) {
            kern = wpe == 2 ? &k_persist<2, true, false, false> : (wpe == 3 ? &k_persist<3, true, false, false>
                                                                             : &k_persist<4, true, false, false>);
        } else {
            kern = wpe == 2 ? &k_persist<2, false, false, false> : (wpe == 3 ? &k_persist<3, false, false, false>
                                                                              : &k_persist<4, false, false, false>);
        }
        lds_bytes = first ? 0 : (size_t)block * sizeof(Owner);
        int per_cu = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds_bytes), "occupancy");
        int64_t resident = (int64_t)std::max(1, per_cu) * s->n_cu;
        if (p->grid_fraction > 0.0f && p->grid_fraction < 1.0f)
            resident = std::max<int64_t>(1, (int64_t)((double)resident * p->grid_fraction));
        grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, resident));
    }
    const int64_t grid0 = (n + 1023) / 1024;
    // zero the counters, optional record padding, k_init, the persistent kernel
    auto enqueue = [&](hipStream_t q, bool events) {
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, q, cnt, 4);
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, q, reinterpret_cast<int32_t*>(ctr), 2 * BBRT_NCOUNTERS);
        if (pad_records && p->record_k > 0) {
            const int64_t total = n * (int64_t)p->record_k;
            const unsigned pg = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
            hipLaunchKernelGGL(k_pad_records, dim3(pg), dim3(256), 0, q, O, total);
        }
        if (group_bits)
            hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, q, hist, (int)nbins);
        hipLaunchKernelGGL(k_init, dim3((unsigned)grid0), dim3(1024), 0, q, d_o, d_d, n, dev, R, O,
                           p->max_bounces, live[0], cnt, ctr, group_bits, bins, hist);
        if (group_bits) {
            hipLaunchKernelGGL(k_scan_bins, dim3(1), dim3(1024), 0, q, hist, cursor, (int)nbins);
            hipLaunchKernelGGL(k_group_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, q, live[0], bins, cnt,
                               cursor, live[1]);
        }
        if (kern) {
            if (events) check(hipEventRecord(s->kev[0], q), "hipEventRecord");
            hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds_bytes, q, dev, R, group_bits ? live[1] : live[0], cnt, cnt + 2, O,
                               p->max_bounces, p->record_k, p->newton_mode, p->flags, ctr, MD, nbatch);
            if (events) check(hipEventRecord(s->kev[1], q), "hipEventRecord");
        }
    };
    if (async) {
        enqueue(st, false);
        if (d_counters)
            hipLaunchKernelGGL(k_counters_out, dim3(1), dim3(64), 0, st, ctr, d_counters, (int64_t)(kern ? 1 : 0));
        check(hipGetLastError(), "bounce kernel launch");
        return;
    }
    int launches = 0;
    bool graph_timed = false;
    if (graph) {
        // One hipGraph per scene handle, captured on the handle's own stream and
        // replayed while every argument is the same.
        GraphKey key{};
        key.d_o = d_o; key.d_d = d_d; key.n = n; key.ray_scene = d_ray_scene; key.params = *p;
        key.outs = O; key.scratch = base; key.grid = grid; key.block = block; key.kern = (const void*)kern;
        key.pad = pad_records;
        hipStream_t gs = s->gstream;
        if (st != gs) {
            check(hipEventRecord(s->ev_dep, st), "hipEventRecord");
            check(hipStreamWaitEvent(gs, s->ev_dep, 0), "hipStreamWaitEvent");
        }
        if (!s->gexec || std::memcmp(&key, &s->gkey, sizeof(GraphKey)) != 0) {
            if (s->gexec) { check(hipGraphExecDestroy(s->gexec), "hipGraphExecDestroy"); s->gexec = nullptr; }
            hipGraph_t g = nullptr;
            check(hipStreamBeginCapture(gs, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
            enqueue(gs, false);
            check(hipStreamEndCapture(gs, &g), "hipStreamEndCapture");
            const hipError_t ie = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            check(ie, "hipGraphInstantiate");
            s->gkey = key;
            s->graph_captures++;
        }
        check(hipEventRecord(s->ev[0], gs), "hipEventRecord");
        check(hipGraphLaunch(s->gexec, gs), "hipGraphLaunch");
        st = gs;
        launches = kern ? 1 : 0;
        graph_timed = true;
    } else if (!wavefront) {
        check(hipEventRecord(s->ev[0], st), "hipEventRecord");
        enqueue(st, timing);
        check(hipGetLastError(), "bounce kernel launch");
        launches = kern ? 1 : 0;
    } else {
        check(hipEventRecord(s->ev[0], st), "hipEventRecord");
        enqueue(st, false);
        check(hipGetLastError(), "k_init launch");
        int cur = 0;
        for (int b = 0; b < p->max_bounces; ++b) {
            check(hipMemcpyAsync(s->h_count, cnt + cur, sizeof(int32_t), hipMemcpyDeviceToHost, st), "count D2H");
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            const int32_t n_live = *s->h_count;
            if (n_live <= 0) break;
            hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, cnt + (cur ^ 1), 1);
            const unsigned gridb = (unsigned)((n_live + block - 1) / block);
            if (timing) check(hipEventRecord(s->kev[2 * launches], st), "hipEventRecord");
            hipLaunchKernelGGL(k_bounce, dim3(gridb), dim3(block), 0, st, dev, R, live[cur], cnt + cur,
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
        if (graph_timed) {
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]), "hipEventElapsedTime");
            ksum = ms;   // the whole graph: zero fills, k_init and the persistent kernel
        } else if (timing)
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

// Runs fn(i) for i in [0, n) on up to `threads` host threads; rethrows the first error.
template <class F>
void for_each_parallel(int32_t n, unsigned threads, F&& fn, const char* who) {
    std::atomic<int32_t> next{ 0 };
    std::vector<std::string> errs((size_t)n);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < std::max(1u, std::min<unsigned>(threads, (unsigned)n)); ++t)
        pool.emplace_back([&]() {
            for (int32_t i = next++; i < n; i = next++) {
                try { fn(i); } catch (const std::exception& e) { errs[i] = e.what(); }
            }
        });
    for (auto& th : pool) th.join();
    for (int32_t i = 0; i < n; ++i)
        if (!errs[i].empty()) throw std::runtime_error(std::string(who) + ": sub-scene " + std::to_string(i) + ": " + errs[i]);
}

// Builds every sub-scene and uploads them as one device scene: patches, coarse
// data and culling boxes concatenated in sub-scene order, each sub-scene's BVH
// appended with its node, leaf and patch indices offset to global ones.
// gpu_build (compact scenes): the host computes only patch records, node-set
// tables, centroids and node boxes; k_patch_stats computes each patch's max
// sagitta, finiteness and longest cell diagonal, the host then inflates the
// boxes and builds the BVHs, and k_patch_pyramid builds the FP32 pyramid.
int create_scenes(const bbrt_patch_set* sets, int32_t n_sets, bbrt_scene* out, const char* who, bool compact,
                  bool gpu_build) {
    *out = nullptr;
    bbrt_scene s = nullptr;
    try {
        int ndev = 0;
        check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (ndev <= 0) throw std::runtime_error(std::string(who) + ": no HIP device");
        if (gpu_build && !compact) throw std::runtime_error(std::string(who) + ": the device build is for compact scenes");
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<bbrt::HostScene> H((size_t)n_sets);
        if (n_sets == 1) bbrt::build_host_scene(sets[0], H[0], (int)hw, compact, gpu_build);
        else for_each_parallel(n_sets, hw, [&](int32_t i) { bbrt::build_host_scene(sets[i], H[i], 1, compact, gpu_build); }, who);
        int64_t n_patch = 0;
        for (const auto& h : H) n_patch += h.n;
        if (n_patch >= (int64_t)1 << 27) throw std::runtime_error(std::string(who) + ": too many patches for one scene handle");
        s = new bbrt_scene_s();
        check(hipGetDevice(&s->device), "hipGetDevice");
        check(hipDeviceGetAttribute(&s->n_cu, hipDeviceAttributeMultiprocessorCount, s->device), "CU count");
        check(hipStreamCreate(&s->stream), "hipStreamCreate");
        check(hipStreamCreateWithFlags(&s->gstream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreateWithFlags(&s->ev_dep, hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreate(&s->ev[0]), "hipEventCreate");
        check(hipEventCreate(&s->ev[1]), "hipEventCreate");
        check(hipHostMalloc(reinterpret_cast<void**>(&s->h_count), sizeof(int32_t) * 64, hipHostMallocDefault),
              "hipHostMalloc");
        s->n_scenes = n_sets;
        s->compact = compact;
        s->patch_begin.assign(1, 0);
        for (const auto& h : H) s->patch_begin.push_back(s->patch_begin.back() + h.n);

        // patch records with global node, field and node-set offsets (fixed before the
        // host arrays are released by the uploads)
        std::vector<int64_t> node_base((size_t)n_sets), field_base((size_t)n_sets);
        std::vector<int32_t> set_base((size_t)n_sets);
        {
            int64_t no = 0, fo = 0;
            int32_t so = 0;
            for (int32_t i = 0; i < n_sets; ++i) {
                node_base[i] = no; field_base[i] = fo; set_base[i] = so;
                if (compact && so + H[i].n_nodesets >= (1 << 16))
                    throw std::runtime_error(std::string(who) + ": more than 65535 distinct node sets");
                no += (int64_t)H[i].nodes.size(); fo += (int64_t)H[i].fields.size(); so += H[i].n_nodesets;
            }
        }
        std::vector<PatchMeta> meta;
        auto merge_meta = [&]() {
            meta.clear();
            meta.reserve((size_t)n_patch);
            for (int32_t i = 0; i < n_sets; ++i) {
                const int32_t so = set_base[i];
                for (PatchMeta m : H[i].meta) {
                    m.node_off += node_base[i]; m.field_off += field_base[i];
                    if (compact) m.nodeset = (((m.nodeset >> 16) + so) << 16) | ((m.nodeset & 0xFFFF) + so);
                    meta.push_back(m);
                }
            }
        };
        // big arrays go straight from each sub-scene to its slice of the device array
        auto upload_parts = [&](auto member) {
            using T = typename std::decay_t<decltype(H[0].*member)>::value_type;
            size_t total = 0;
            for (const auto& h : H) total += (h.*member).size();
            T* d = nullptr;
            check(hipMalloc(&d, std::max<size_t>(1, total) * sizeof(T)), "hipMalloc(scene)");
            s->allocs.push_back(d);
            s->bytes += (int64_t)(std::max<size_t>(1, total) * sizeof(T));
            size_t at = 0;
            for (auto& h : H) {
                auto& v = h.*member;
                if (!v.empty()) check(hipMemcpy(d + at, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(scene)");
                at += v.size();
                std::vector<T>().swap(v);   // release host memory as it goes
            }
            return d;
        };
        DevScene& D = s->dev;
        merge_meta();
        PatchMeta* d_meta = const_cast<PatchMeta*>(dev_upload(meta, s->allocs, s->bytes));
        D.meta = d_meta;
        D.ff_rows = compact ? upload_parts(&bbrt::HostScene::ff_rows) : nullptr;
        D.mid_basis = compact ? upload_parts(&bbrt::HostScene::mid_basis) : nullptr;
        D.compact = compact ? 1 : 0;
        D.nodes = upload_parts(&bbrt::HostScene::nodes);
        D.fields = upload_parts(&bbrt::HostScene::fields);
        D.cxyz = upload_parts(&bbrt::HostScene::cxyz);
        D.cell = upload_parts(&bbrt::HostScene::cell);
        D.box_cell = upload_parts(&bbrt::HostScene::box_cell);
        D.ray_scene = nullptr;
        if (gpu_build) {
            // pass 1 on the device, then the host side of every sub-scene
            std::vector<double> ms((size_t)n_patch), hm((size_t)n_patch);
            std::vector<int32_t> fin((size_t)n_patch);
            double* d_ms = nullptr;
            double* d_hm = nullptr;
            int32_t* d_fin = nullptr;
            check(hipMalloc(&d_ms, n_patch * sizeof(double)), "hipMalloc");
            check(hipMalloc(&d_hm, n_patch * sizeof(double)), "hipMalloc");
            check(hipMalloc(&d_fin, n_patch * sizeof(int32_t)), "hipMalloc");
            hipLaunchKernelGGL(k_patch_stats, dim3((unsigned)n_patch), dim3(256), 0, s->stream, D, d_ms, d_fin, d_hm);
            check(hipGetLastError(), "k_patch_stats");
            check(hipMemcpyAsync(ms.data(), d_ms, n_patch * sizeof(double), hipMemcpyDeviceToHost, s->stream), "D2H");
            check(hipMemcpyAsync(hm.data(), d_hm, n_patch * sizeof(double), hipMemcpyDeviceToHost, s->stream), "D2H");
            check(hipMemcpyAsync(fin.data(), d_fin, n_patch * sizeof(int32_t), hipMemcpyDeviceToHost, s->stream), "D2H");
            check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
            (void)hipFree(d_ms); (void)hipFree(d_hm); (void)hipFree(d_fin);
            for_each_parallel(n_sets, hw, [&](int32_t i) {
                const int64_t p0 = s->patch_begin[i], pn = H[i].n;
                bbrt::finish_host_scene(H[i], std::vector<double>(ms.begin() + p0, ms.begin() + p0 + pn),
                                        std::vector<int32_t>(fin.begin() + p0, fin.begin() + p0 + pn),
                                        std::vector<double>(hm.begin() + p0, hm.begin() + p0 + pn), false);
            }, who);
            merge_meta();
            check(hipMemcpy(d_meta, meta.data(), meta.size() * sizeof(PatchMeta), hipMemcpyHostToDevice), "H2D meta");
        }
        // BVHs, leaves and the sub-scene table
        std::vector<BvhNode> bvh;
        std::vector<int32_t> leaf;
        if (gpu_build) {
            // every sub-scene's BVH built on the device from the inflated boxes and centroids
            std::vector<double> gbox, gcent, widen;
            gbox.reserve((size_t)n_patch * 6);
            gcent.reserve((size_t)n_patch * 3);
            for (const auto& h : H) {
                gbox.insert(gbox.end(), h.pbox.begin(), h.pbox.end());
                gcent.insert(gcent.end(), h.pcent.begin(), h.pcent.end());
                widen.push_back(h.cull_widen);
            }
            std::vector<void*> tmp;
            int64_t tmp_bytes = 0;
            bbrt::DeviceTree T;
            try {
                const double* d_gbox = dev_upload(gbox, tmp, tmp_bytes);
                const double* d_gcent = dev_upload(gcent, tmp, tmp_bytes);
                bbrt::build_bvh_device(d_gbox, d_gcent, s->patch_begin, 4, widen.data(), s->stream, s->allocs, s->bytes,
                                       T);
            } catch (...) {
                for (void* p : tmp) (void)hipFree(p);
                throw;
            }
            for (void* p : tmp) (void)hipFree(p);
            D.bvh = T.bvh;
            D.bvh4 = T.bvh4;
            D.leaf = T.leaf;
            D.leaf_box = T.leaf_box;
            bvh.resize((size_t)T.n_nodes);
            leaf.resize((size_t)n_patch);
            check(hipMemcpy(bvh.data(), T.bvh, bvh.size() * sizeof(BvhNode), hipMemcpyDeviceToHost), "D2H bvh");
            check(hipMemcpy(leaf.data(), T.leaf, leaf.size() * sizeof(int32_t), hipMemcpyDeviceToHost), "D2H leaf");
            for (int32_t i = 0; i < n_sets; ++i) {
                bbrt::SceneInfo si{};
                si.root = T.roots[i]; si.patch_begin = (int32_t)s->patch_begin[i]; si.diag = H[i].diag;
                si.cull_limit = 8.0 * H[i].scale; si.cull_widen = H[i].cull_widen;
                s->scenes.push_back(si);
            }
            if ((int64_t)bvh.size() >= (int64_t)1 << 28) throw std::runtime_error(std::string(who) + ": too many BVH nodes");
        } else {
            for (int32_t i = 0; i < n_sets; ++i) {
                const auto& h = H[i];
                const int32_t bo = (int32_t)bvh.size(), lo = (int32_t)leaf.size();
                for (BvhNode nd : h.bvh) {
                    if (nd.right >= 0) nd.right += bo; else nd.start += lo;
                    bvh.push_back(nd);
                }
                for (int32_t o : h.leaf) leaf.push_back(o + (int32_t)s->patch_begin[i]);
                bbrt::SceneInfo si{};
                si.root = bo; si.patch_begin = (int32_t)s->patch_begin[i]; si.diag = h.diag; si.cull_limit = 8.0 * h.scale;
                si.cull_widen = h.cull_widen;
                s->scenes.push_back(si);
            }
            if ((int64_t)bvh.size() >= (int64_t)1 << 28) throw std::runtime_error(std::string(who) + ": too many BVH nodes");
            for (const BvhNode& nd : bvh)
                if (nd.right >= (1 << 28) || nd.start >= (1 << 27) || nd.count > 7)
                    throw std::runtime_error(std::string(who) + ": BVH exceeds the packed node limits");
            D.bvh = dev_upload(bvh, s->allocs, s->bytes);
            {
                std::vector<int32_t> roots;
                std::vector<double> widen;
                for (const auto& si : s->scenes) { roots.push_back(si.root); widen.push_back(si.cull_widen); }
                std::vector<bbrt::Bvh4Node> bvh4;
                bbrt::build_bvh4(bvh, roots, widen, bvh4);
                D.bvh4 = dev_upload(bvh4, s->allocs, s->bytes);
            }
            D.leaf = dev_upload(leaf, s->allocs, s->bytes);
            {
                std::vector<double> lb(6 * std::max<size_t>(1, leaf.size()), 0.0);
                for (const BvhNode& nd : bvh)
                    if (nd.right < 0)
                        for (int a = 0; a < 3; ++a) { lb[6 * (size_t)nd.start + a] = nd.lo[a]; lb[6 * (size_t)nd.start + 3 + a] = nd.hi[a]; }
                D.leaf_box = dev_upload(lb, s->allocs, s->bytes);
            }
        }
        s->d_scenes = const_cast<bbrt::SceneInfo*>(dev_upload(s->scenes, s->allocs, s->bytes));
        D.scenes = s->d_scenes;
        if (gpu_build) {
            auto dalloc = [&](size_t floats) {
                float* d = nullptr;
                check(hipMalloc(&d, std::max<size_t>(1, floats) * sizeof(float)), "hipMalloc(scene)");
                s->allocs.push_back(d);
                s->bytes += (int64_t)(std::max<size_t>(1, floats) * sizeof(float));
                return d;
            };
            float* bt = dalloc((size_t)n_patch * bbrt::kTileFloats);
            float* bk = dalloc((size_t)n_patch * bbrt::kBlockFloats);
            float* bs = dalloc((size_t)n_patch * bbrt::kSuperFloats);
            float* bp = dalloc((size_t)n_patch * 8);
            uint8_t* qb = nullptr;
            check(hipMalloc(&qb, std::max<size_t>(1, (size_t)n_patch * bbrt::kQCellPatchBytes)), "hipMalloc(scene)");
            s->allocs.push_back(qb);
            s->bytes += (int64_t)n_patch * (int64_t)bbrt::kQCellPatchBytes;
            // (the quantized boxes of a patch without culling are never read and stay unwritten)
            for (int32_t i = 0; i < n_sets; ++i)
                if (H[i].n > 0)
                    hipLaunchKernelGGL(k_patch_pyramid, dim3((unsigned)H[i].n), dim3(256), 0, s->stream, D,
                                       s->patch_begin[i], H[i].cull_widen, bt, bk, bs, bp, qb);
            D.qbox_cell = qb;
            check(hipGetLastError(), "k_patch_pyramid");
            check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
            D.box_tile = bt; D.box_block = bk; D.box_super = bs; D.box_patch = bp;
        } else {
            D.qbox_cell = compact ? upload_parts(&bbrt::HostScene::qbox_cell) : nullptr;
            D.box_tile = upload_parts(&bbrt::HostScene::box_tile);
            D.box_block = upload_parts(&bbrt::HostScene::box_block);
            D.box_super = upload_parts(&bbrt::HostScene::box_super);
            D.box_patch = upload_parts(&bbrt::HostScene::box_patch);
        }
        {
            float* lp = nullptr;
            const size_t nl = std::max<size_t>(1, leaf.size());
            check(hipMalloc(&lp, nl * 8 * sizeof(float)), "hipMalloc(scene)");
            s->allocs.push_back(lp);
            s->bytes += (int64_t)(nl * 8 * sizeof(float));
            if (!leaf.empty()) {
                const int64_t total = (int64_t)leaf.size() * 2;
                hipLaunchKernelGGL(k_gather_leaf_pbox, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s->stream,
                                   D.leaf, D.box_patch, (int64_t)leaf.size(), lp);
                check(hipGetLastError(), "k_gather_leaf_pbox");
                check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
            }
            D.leaf_pbox = lp;
        }
        s->n_patches = n_patch;
        for (auto& h : H) {
            s->pbox.insert(s->pbox.end(), h.pbox.begin(), h.pbox.end());
            s->pcent.insert(s->pcent.end(), h.pcent.begin(), h.pcent.end());
            for (int64_t k = 0; k < h.n; ++k) s->max_sag.push_back(h.meta[k].max_sag);
        }
        s->bvh = std::move(bvh);
        s->leaf = std::move(leaf);
        s->diag = H[0].diag;
        s->max_patch_sag = 0.0;
        for (const auto& h : H) s->max_patch_sag = std::max(s->max_patch_sag, h.max_patch_sag);
        *out = s;
        return 0;
    } catch (const std::exception& e) {
        delete s;
        return fail(e);
    }
}

// Device per-ray sub-scene ids for the ray ranges `begin` (cached while they do not change).
const int32_t* ray_scene_ids(bbrt_scene s, const int64_t* begin, hipStream_t st) {
    const int64_t n = begin[s->n_scenes];
    std::vector<int64_t> key(begin, begin + s->n_scenes + 1);
    if (key == s->ray_scene_key && s->d_ray_scene) return s->d_ray_scene;
    // traces still in flight on any stream may read the current ids
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (n > s->ray_scene_cap) {
        if (s->d_ray_scene) check(hipFree(s->d_ray_scene), "hipFree");
        s->d_ray_scene = nullptr;
        s->ray_scene_cap = 0;
        check(hipMalloc(&s->d_ray_scene, std::max<int64_t>(1, n) * sizeof(int32_t)), "hipMalloc(ray scene ids)");
        s->ray_scene_cap = n;
    }
    int64_t* d_begin = nullptr;
    check(hipMalloc(&d_begin, key.size() * sizeof(int64_t)), "hipMalloc");
    check(hipMemcpyAsync(d_begin, key.data(), key.size() * sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
    if (n > 0)
        hipLaunchKernelGGL(k_ray_scene, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_begin, s->n_scenes, n,
                           s->d_ray_scene);
    check(hipGetLastError(), "k_ray_scene");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    check(hipFree(d_begin), "hipFree");
    s->ray_scene_key = std::move(key);
    return s->d_ray_scene;
}

void check_ranges(bbrt_scene s, const int64_t* begin) {
    if (!begin) throw std::runtime_error("bbrt_trace_batch: ray_begin is NULL");
    if (begin[0] != 0) throw std::runtime_error("bbrt_trace_batch: ray_begin[0] must be 0");
    for (int32_t i = 0; i < s->n_scenes; ++i)
        if (begin[i + 1] < begin[i]) throw std::runtime_error("bbrt_trace_batch: ray_begin must not decrease");
}

int trace_host(bbrt_scene s, const double* origins, const double* dirs, int64_t n, const int64_t* ray_begin,
               const bbrt_params* p, const bbrt_outputs* out, int64_t* counters, const char* who);

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
    return create_scenes(ps, 1, out, "bbrt_scene_create", false, false);
}

int bbrt_scene_create_batch(const bbrt_patch_set* sets, int32_t n_sets, bbrt_scene* out) {
    if (!sets || !out) { g_last_error = "bbrt_scene_create_batch: NULL argument"; return -1; }
    if (n_sets <= 0) { g_last_error = "bbrt_scene_create_batch: no sub-scenes"; return -1; }
    return create_scenes(sets, n_sets, out, "bbrt_scene_create_batch", false, false);
}

int bbrt_scene_create_ex(const bbrt_patch_set* sets, int32_t n_sets, uint32_t scene_flags, bbrt_scene* out) {
    if (!sets || !out) { g_last_error = "bbrt_scene_create_ex: NULL argument"; return -1; }
    if (n_sets <= 0) { g_last_error = "bbrt_scene_create_ex: no sub-scenes"; return -1; }
    if (scene_flags & ~(uint32_t)(BBRT_SCENE_COMPACT | BBRT_SCENE_DEVICE_BUILD)) {
        g_last_error = "bbrt_scene_create_ex: unknown flag";
        return -1;
    }
    if ((scene_flags & BBRT_SCENE_DEVICE_BUILD) && !(scene_flags & BBRT_SCENE_COMPACT)) {
        g_last_error = "bbrt_scene_create_ex: BBRT_SCENE_DEVICE_BUILD needs BBRT_SCENE_COMPACT";
        return -1;
    }
    return create_scenes(sets, n_sets, out, "bbrt_scene_create_ex", (scene_flags & BBRT_SCENE_COMPACT) != 0,
                         (scene_flags & BBRT_SCENE_DEVICE_BUILD) != 0);
}

int bbrt_scene_batch_info(bbrt_scene s, int32_t* n_scenes, int64_t* patch_begin, int32_t* bvh_root, double* diag) {
    if (!s) { g_last_error = "bbrt_scene_batch_info: NULL scene"; return -1; }
    if (n_scenes) *n_scenes = s->n_scenes;
    for (int32_t i = 0; i < s->n_scenes; ++i) {
        if (bvh_root) bvh_root[i] = s->scenes[i].root;
        if (diag) diag[i] = s->scenes[i].diag;
    }
    if (patch_begin) std::memcpy(patch_begin, s->patch_begin.data(), s->patch_begin.size() * sizeof(int64_t));
    return 0;
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
    if (s->compact) { g_last_error = "bbrt_scene_coarse: a compact scene keeps no coarse grid"; return -1; }
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
        trace_impl(s, origins, dirs, n, p, out, counters, static_cast<hipStream_t>(stream), kernel_ms, false, nullptr);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_trace_device_async(bbrt_scene s, const double* origins, const double* dirs, int64_t n,
                            const int64_t* ray_begin, const bbrt_params* p, const bbrt_outputs* out,
                            int64_t* d_counters, void* stream) {
    if (!s) { g_last_error = "bbrt_trace_device_async: NULL scene"; return -1; }
    try {
        std::lock_guard<std::mutex> lock(s->async_mu);
        check(hipSetDevice(s->device), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
        const int32_t* ids = nullptr;
        if (ray_begin) {
            check_ranges(s, ray_begin);
            if (ray_begin[s->n_scenes] != n)
                throw std::runtime_error("bbrt_trace_device_async: n_rays must equal ray_begin[n_scenes]");
            ids = ray_scene_ids(s, ray_begin, st);
        }
        trace_impl(s, origins, dirs, n, p, out, nullptr, st, nullptr, false, ids, true, d_counters);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_trace_device_noalloc(bbrt_scene s, const double* origins, const double* dirs, int64_t n,
                              const bbrt_params* p, const bbrt_outputs* out, int64_t* d_counters, void* stream) {
    if (!s) { g_last_error = "bbrt_trace_device_noalloc: NULL scene"; return -1; }
    try {
        if (s->n_scenes > 1)
            throw std::runtime_error("bbrt_trace_device_noalloc: one-scene handles only (a batch handle's per-ray "
                                     "sub-scene ids would need device memory)");
        std::lock_guard<std::mutex> lock(s->async_mu);
        check(hipSetDevice(s->device), "hipSetDevice");
        char* base = noalloc_scratch(s->device);
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
        trace_impl(s, origins, dirs, n, p, out, nullptr, st, nullptr, false, nullptr, true, d_counters, base);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_bounce_map(bbrt_scene s, const int64_t* patches, int64_t n_sample, const bbrt_bounce_map_params* bp,
                    const bbrt_bounce_map_outputs* out, int64_t* counters) {
    if (!s || !bp || (n_sample > 0 && !patches)) { g_last_error = "bbrt_bounce_map: NULL argument"; return -1; }
    try {
        if (s->n_scenes > 1) throw std::runtime_error("bbrt_bounce_map: one-scene handles only");
        if (bp->grid_n < 2) throw std::runtime_error("bbrt_bounce_map: grid_n must be >= 2");
        if (bp->theta_n < 1 || bp->phi_n < 1) throw std::runtime_error("bbrt_bounce_map: theta_n and phi_n must be >= 1");
        if (!bp->cos_phi || !bp->sin_phi || !bp->eps) throw std::runtime_error("bbrt_bounce_map: NULL table");
        if (n_sample < 0) throw std::runtime_error("bbrt_bounce_map: negative patch count");
        for (int64_t q = 0; q < n_sample; ++q)
            if (patches[q] < 0 || patches[q] >= s->n_patches)
                throw std::runtime_error("bbrt_bounce_map: patch index out of range");
        const int64_t nn = (int64_t)bp->grid_n * bp->grid_n;
        const int64_t n_points = n_sample * nn;
        const int per = bp->theta_n * bp->phi_n;
        const int64_t n_rays = n_points * per;
        if (n_rays > (int64_t)INT32_MAX - 1024) throw std::runtime_error("bbrt_bounce_map: too many launch rays");
        if (n_points == 0) {
            if (counters) std::memset(counters, 0, BBRT_NCOUNTERS * sizeof(int64_t));
            return 0;
        }
        check(hipSetDevice(s->device), "hipSetDevice");
        CallBuffers mem;
        hipStream_t st = s->stream;
        int64_t* d_pid = mem.get<int64_t>(n_sample);
        double* d_eps = mem.get<double>(n_sample);
        double* d_cp = mem.get<double>(bp->phi_n);
        double* d_sp = mem.get<double>(bp->phi_n);
        check(hipMemcpyAsync(d_pid, patches, n_sample * sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
        check(hipMemcpyAsync(d_eps, bp->eps, n_sample * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
        check(hipMemcpyAsync(d_cp, bp->cos_phi, bp->phi_n * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
        check(hipMemcpyAsync(d_sp, bp->sin_phi, bp->phi_n * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
        double* d_grid = (out && out->grid) ? mem.get<double>(9 * (size_t)n_points) : nullptr;
        double* d_o = mem.get<double>(3 * (size_t)n_rays);
        double* d_d = mem.get<double>(3 * (size_t)n_rays);
        int8_t* d_st = mem.get<int8_t>(n_rays);
        int32_t* d_bc = mem.get<int32_t>(n_rays);
        DevScene dev = s->dev;
        dev.ray_scene = nullptr;
        hipLaunchKernelGGL(k_bmap_launch, dim3((unsigned)((n_points + 255) / 256)), dim3(256), 0, st, dev, d_pid, d_eps,
                           n_sample, bp->grid_n, bp->theta_n, bp->phi_n, bp->cos_theta_max, d_cp, d_sp, d_grid, d_o, d_d);
        check(hipGetLastError(), "k_bmap_launch");
        bbrt_params p{};
        p.newton_mode = bp->newton_mode;
        p.seed_mode = BBRT_SEED_TRIANGLE;
        p.max_bounces = bp->max_bounces;
        p.record_k = 0;
        p.flags = bp->flags & ~(uint32_t)(BBRT_FLAG_GRAPH | BBRT_FLAG_TIME_KERNELS);
        bbrt_outputs to{};
        to.status = d_st;
        to.bounces = d_bc;
        trace_impl(s, d_o, d_d, n_rays, &p, &to, counters, st, nullptr, false, nullptr);
        int32_t* d_mb = mem.get<int32_t>(n_points);
        int32_t* d_ec = mem.get<int32_t>(n_points);
        int32_t* d_bd = mem.get<int32_t>(n_points);
        double* d_bod = (out && out->best_od) ? mem.get<double>(6 * (size_t)n_points) : nullptr;
        hipLaunchKernelGGL(k_bmap_reduce, dim3((unsigned)((n_points + 255) / 256)), dim3(256), 0, st, d_st, d_bc, d_o, d_d,
                           n_points, per, d_mb, d_ec, d_bd, d_bod);
        check(hipGetLastError(), "k_bmap_reduce");
        auto down = [&](void* h, const void* dsrc, size_t bytes) {
            if (h) check(hipMemcpyAsync(h, dsrc, bytes, hipMemcpyDeviceToHost, st), "D2H");
        };
        if (out) {
            down(out->grid, d_grid, 9 * (size_t)n_points * sizeof(double));
            down(out->max_bounces, d_mb, (size_t)n_points * sizeof(int32_t));
            down(out->error_count, d_ec, (size_t)n_points * sizeof(int32_t));
            down(out->best_dir, d_bd, (size_t)n_points * sizeof(int32_t));
            down(out->best_od, d_bod, 6 * (size_t)n_points * sizeof(double));
            down(out->launch_o, d_o, 3 * (size_t)n_rays * sizeof(double));
            down(out->launch_d, d_d, 3 * (size_t)n_rays * sizeof(double));
        }
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_trace(bbrt_scene s, const double* origins, const double* dirs, int64_t n, const bbrt_params* p,
               const bbrt_outputs* out, int64_t* counters) {
    return trace_host(s, origins, dirs, n, nullptr, p, out, counters, "bbrt_trace");
}

int bbrt_trace_batch(bbrt_scene s, const double* origins, const double* dirs, const int64_t* ray_begin,
                     const bbrt_params* p, const bbrt_outputs* out, int64_t* counters) {
    if (!s || !ray_begin) { g_last_error = "bbrt_trace_batch: NULL scene or ray_begin"; return -1; }
    return trace_host(s, origins, dirs, ray_begin[s->n_scenes], ray_begin, p, out, counters, "bbrt_trace_batch");
}

int bbrt_trace_batch_device(bbrt_scene s, const double* origins, const double* dirs, const int64_t* ray_begin,
                            const bbrt_params* p, const bbrt_outputs* out, int64_t* counters, void* stream,
                            double* kernel_ms) {
    if (!s) { g_last_error = "bbrt_trace_batch_device: NULL scene"; return -1; }
    try {
        check_ranges(s, ray_begin);
        check(hipSetDevice(s->device), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
        const int32_t* ids = ray_scene_ids(s, ray_begin, st);
        trace_impl(s, origins, dirs, ray_begin[s->n_scenes], p, out, counters, st, kernel_ms, false, ids);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int bbrt_autotune(bbrt_scene s, const double* origins, const double* dirs, const int64_t* ray_begin, int64_t n,
                  const bbrt_params* base, const bbrt_outputs* out, bbrt_params* best, bbrt_tune_entry* table,
                  int32_t max_entries, int32_t* n_entries, void* stream) {
    if (!s || !base || !best) { g_last_error = "bbrt_autotune: NULL argument"; return -1; }
    try {
        check(hipSetDevice(s->device), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
        const int32_t* ids = nullptr;
        if (ray_begin) {
            check_ranges(s, ray_begin);
            ids = ray_scene_ids(s, ray_begin, st);
            n = ray_begin[s->n_scenes];
        }
        if (base->flags & BBRT_FLAG_WAVEFRONT) throw std::runtime_error("bbrt_autotune: tunes the persistent schedule only");
        int32_t count = 0;
        auto run = [&](const bbrt_params& q) {
            double best_ms = 1e300;
            for (int rep = 0; rep < 2; ++rep) {
                double kms[3] = { 0, 0, 0 };
                trace_impl(s, origins, dirs, n, &q, out, nullptr, st, kms, false, ids);
                best_ms = std::min(best_ms, kms[2]);
            }
            if (table && count < max_entries)
                table[count] = bbrt_tune_entry{ q.block, q.waves_per_eu, q.newton_batch, q.grid_fraction, best_ms };
            ++count;
            return best_ms;
        };
        // coordinate descent from the library defaults over occupancy, Newton batch, block size and grid size
        bbrt_params cur = *base;
        cur.flags &= ~(uint32_t)BBRT_FLAG_GRAPH;
        cur.block = default_block(base->seed_mode == BBRT_SEED_LEARNED, s->compact != 0);
        cur.waves_per_eu = default_waves_per_eu(s->compact); cur.newton_batch = default_newton_batch(s->compact != 0);
        cur.grid_fraction = 1.0f;
        double cur_ms = run(cur);
        auto sweep = [&](auto setter, std::initializer_list<double> values) {
            bbrt_params keep = cur;
            for (double v : values) {
                bbrt_params q = cur;
                setter(q, v);
                if (std::memcmp(&q, &cur, sizeof(q)) == 0) continue;
                const double ms = run(q);
                if (ms < cur_ms) { cur_ms = ms; keep = q; }
            }
            cur = keep;
        };
        sweep([](bbrt_params& q, double v) { q.waves_per_eu = (int)v; }, { 2, 3, 4 });
        sweep([](bbrt_params& q, double v) { q.newton_batch = (int)v; }, { 8, 16, 32, 64 });
        sweep([](bbrt_params& q, double v) { q.block = (int)v; }, { 64, 128, 256 });
        sweep([](bbrt_params& q, double v) { q.grid_fraction = (float)v; }, { 0.5, 0.75, 1.0 });
        *best = cur;
        best->flags = base->flags;
        if (n_entries) *n_entries = count;
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

}  // extern "C"

namespace {

int trace_host(bbrt_scene s, const double* origins, const double* dirs, int64_t n, const int64_t* ray_begin,
               const bbrt_params* p, const bbrt_outputs* out, int64_t* counters, const char* who) {
    if (!s) { g_last_error = std::string(who) + ": NULL scene"; return -1; }
    if (!p) { g_last_error = std::string(who) + ": params is NULL"; return -1; }
    if (n > 0 && (!origins || !dirs)) { g_last_error = std::string(who) + ": NULL origins or dirs"; return -1; }
    std::vector<void*> tmp;
    auto cleanup = [&]() { for (void* q : tmp) (void)hipFree(q); };
    try {
        check(hipSetDevice(s->device), "hipSetDevice");
        const int32_t* ids = nullptr;
        if (ray_begin) {
            check_ranges(s, ray_begin);
            ids = ray_scene_ids(s, ray_begin, s->stream);
        }
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
        if (hout.seed_stats) dout.seed_stats = static_cast<uint32_t*>(dalloc(nn * 8));
        if (hout.hit_count) {
            dout.hit_count = static_cast<int64_t*>(dalloc((size_t)s->n_patches * 8));
            check(hipMemcpy(dout.hit_count, hout.hit_count, (size_t)s->n_patches * 8, hipMemcpyHostToDevice), "H2D hit_count");
        }
        trace_impl(s, d_o, d_d, n, p, &dout, counters, s->stream, nullptr, true, ids);
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
        back(hout.seed_stats, dout.seed_stats, nn * 8);
        cleanup();
        return 0;
    } catch (const std::exception& e) {
        cleanup();
        return fail(e);
    }
}

}  // namespace
