// bbrt_bvh.hip -- BVH::build (include/BVH.hpp:136-189) on the device.
//
// The reference builds its tree recursively: a node's box is the union of its
// objects' boxes in their current order, a node of more than leaf_threshold
// objects splits them with std::nth_element at the middle along the axis of
// largest centroid spread, the left half becomes node index + 1 and the right
// half follows the left subtree (depth-first layout), and leaves append their
// objects in order.  Two facts make that a level-by-level device build:
//   * the node's halves are contiguous ranges of one working order, so the
//     leaf array is that order once every split has run, and a leaf starts
//     where its range starts;
//   * the tree's shape depends only on the object count (halves of n/2 and
//     n - n/2), so every node's index, range and right-child index are known
//     before any key is compared.
// The host lays out those nodes (BvhTask) depth by depth; one workgroup per
// node then folds the boxes and centroid bounds, picks the axis, and runs the
// library's nth_element order on the node's range (bvh_select.h): in LDS by one
// lane when the range fits, else round by round on device memory with the
// whole workgroup (each partition round computed from prefix-summed ranks)
// until the live range fits, then in LDS.  The folds reproduce the reference's sequential min/max exactly
// (first occurrence among equal values, a NaN start kept), so boxes match bit
// for bit, signed zeros included.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "runtime.h"
#include "bvh_device.h"
#include "bvh_select.h"

namespace {

using bbrt::sel::KeyId;
using bbrt::sel::KeyLess;
using bbrt::sel::SelectState;

// One node of the build: its objects are positions begin .. begin+n-1 of the working order.
struct BvhTask {
    int32_t begin, n;
    int32_t node;    // its index in the forest
    int32_t right;   // its right child's index; -1 for a leaf
    int32_t depth;   // below its part's root
    int32_t part;
};

constexpr int kLdsCap = 3072;   // largest live select range staged in LDS (16 bytes an element, 48 KB)

// lo = lesser(lo, x) over a node's objects in order (BVH.hpp:141-143, :160-166) as a reduction of
// (value, position) pairs: numbers before NaN, the smaller value first, equal values (+0 and -0 too) by
// position.  The sequential fold's result is the winner's value, or the first element's when that is NaN.
__device__ __forceinline__ void fold_lo(double& v, int32_t& i, double v2, int32_t i2) {
    const bool n1 = v != v, n2 = v2 != v2;
    const bool take = (n1 != n2) ? n1 : (n1 ? i2 < i : (v2 < v || (!(v < v2) && i2 < i)));
    if (take) { v = v2; i = i2; }
}
__device__ __forceinline__ void fold_hi(double& v, int32_t& i, double v2, int32_t i2) {
    const bool n1 = v != v, n2 = v2 != v2;
    const bool take = (n1 != n2) ? n1 : (n1 ? i2 < i : (v < v2 || (!(v2 < v) && i2 < i)));
    if (take) { v = v2; i = i2; }
}

constexpr int kMaxWaves = 16;   // workgroups of up to 1024 threads

struct BlockShared {
    double v[kMaxWaves][6];
    int32_t i[kMaxWaves][6];
    int32_t part[2][kMaxWaves];
    int32_t bcast[4];
    SelectState st;
    int32_t heap_done;
};

// Folds three lo and three hi coordinates of `vals` (stride `stride`, lo at +0..2, hi at +hi_off..)
// over the objects perm[begin .. begin+n) across the workgroup; every thread gets the result.
__device__ __forceinline__ void fold_bounds(const double* __restrict__ vals, int stride, int hi_off,
                                            const int32_t* __restrict__ perm, int32_t begin, int32_t n,
                                            BlockShared& sh, double lo[3], double hi[3]) {
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x, lane = tid & 63, w = tid >> 6, nw = nt >> 6;
    int32_t ilo[3], ihi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = hi[a] = __builtin_nan(""); ilo[a] = ihi[a] = INT_MAX; }
    for (int32_t j = tid; j < n; j += nt) {
        const double* x = vals + (size_t)stride * perm[begin + j];
        for (int a = 0; a < 3; ++a) { fold_lo(lo[a], ilo[a], x[a], j); fold_hi(hi[a], ihi[a], x[hi_off + a], j); }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a) {
            const double l2 = __shfl_xor(lo[a], off), h2 = __shfl_xor(hi[a], off);
            const int32_t il = __shfl_xor(ilo[a], off), ih = __shfl_xor(ihi[a], off);
            fold_lo(lo[a], ilo[a], l2, il);
            fold_hi(hi[a], ihi[a], h2, ih);
        }
    if (nw > 1) {
        if (lane == 0)
            for (int a = 0; a < 3; ++a) {
                sh.v[w][a] = lo[a]; sh.i[w][a] = ilo[a]; sh.v[w][3 + a] = hi[a]; sh.i[w][3 + a] = ihi[a];
            }
        __syncthreads();
        for (int u = 0; u < nw; ++u)
            for (int a = 0; a < 3; ++a) {
                fold_lo(lo[a], ilo[a], sh.v[u][a], sh.i[u][a]);
                fold_hi(hi[a], ihi[a], sh.v[u][3 + a], sh.i[u][3 + a]);
            }
        __syncthreads();
    }
    const double* x0 = vals + (size_t)stride * perm[begin];
    for (int a = 0; a < 3; ++a) {
        if (x0[a] != x0[a]) lo[a] = x0[a];
        if (x0[hi_off + a] != x0[hi_off + a]) hi[a] = x0[hi_off + a];
    }
}

// Exclusive prefix sums of two counts across the workgroup; tot gets their totals.
__device__ __forceinline__ void block_scan2(int32_t c0, int32_t c1, BlockShared& sh, int32_t& e0, int32_t& e1,
                                            int32_t& t0, int32_t& t1) {
    const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6, nw = (int)blockDim.x >> 6;
    int32_t x0 = c0, x1 = c1;
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y0 = __shfl_up(x0, off), y1 = __shfl_up(x1, off);
        if (lane >= off) { x0 += y0; x1 += y1; }
    }
    if (lane == 63) { sh.part[0][w] = x0; sh.part[1][w] = x1; }
    __syncthreads();
    int32_t b0 = 0, b1 = 0;
    t0 = 0; t1 = 0;
    for (int u = 0; u < nw; ++u) {
        if (u < w) { b0 += sh.part[0][u]; b1 += sh.part[1][u]; }
        t0 += sh.part[0][u]; t1 += sh.part[1][u];
    }
    __syncthreads();
    e0 = b0 + x0 - c0;
    e1 = b1 + x1 - c1;
}

// One partition round of introselect over a[lo+1 .. hi) around the pivot a[lo], by the whole workgroup,
// in the rank form of bvh_select.h's partition_by_ranks: the left scan's stoppers (not below the pivot) go to
// lpos by rank from the left, the right scan's (not above it, the pivot included) to rpos by rank from the
// right, the first K pairs with lpos[k] < rpos[k] are swapped, and the round ends at min(lpos[K], rpos[K-1]).
// Every thread gets the cut; the array is complete on return.
__device__ int32_t partition_round(KeyId* __restrict__ a, int32_t lo, int32_t hi, int32_t* __restrict__ lpos,
                                   int32_t* __restrict__ rpos, BlockShared& sh) {
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const double P = a[lo].key;
    int32_t cl = 0, cr = 0;
    for (int32_t p = lo + tid; p < hi; p += nt) {
        const double x = a[p].key;
        cl += (p > lo && !(x < P)) ? 1 : 0;
        cr += !(P < x) ? 1 : 0;
    }
    int32_t d0, d1, nl, nr;
    block_scan2(cl, cr, sh, d0, d1, nl, nr);
    int32_t carry_l = 0, carry_r = 0, kcount = 0;
    constexpr int kItems = 8;
    for (int32_t cb = lo; cb < hi; cb += nt * kItems) {
        const int32_t p0 = cb + tid * kItems;
        uint32_t fl = 0, fr = 0;
        for (int it = 0; it < kItems; ++it) {
            const int32_t p = p0 + it;
            if (p < hi) {
                const double x = a[p].key;
                if (p > lo && !(x < P)) fl |= 1u << it;
                if (!(P < x)) fr |= 1u << it;
            }
        }
        int32_t el, er, tl, tr;
        block_scan2(__popc(fl), __popc(fr), sh, el, er, tl, tr);
        int32_t rank_l = carry_l + el, r_incl = carry_r + er;
        for (int it = 0; it < kItems; ++it) {
            const int32_t p = p0 + it;
            if ((fr >> it) & 1u) {
                ++r_incl;
                rpos[nr - r_incl] = p;
            }
            if ((fl >> it) & 1u) {
                lpos[rank_l] = p;
                if (nr - r_incl > rank_l) ++kcount;
                ++rank_l;
            }
        }
        carry_l += tl;
        carry_r += tr;
    }
    int32_t dk, dz, K, zz;
    block_scan2(kcount, 0, sh, dk, dz, K, zz);   // also orders the lpos / rpos stores before the swaps
    for (int32_t k = tid; k < K; k += nt) {
        const int32_t x = lpos[k], y = rpos[k];
        const KeyId t = a[x];
        a[x] = a[y];
        a[y] = t;
    }
    if (tid == 0) {
        int32_t cut;
        if (K == 0) cut = nl > 0 ? lpos[0] : hi - 1;
        else cut = (K < nl && lpos[K] < rpos[K - 1]) ? lpos[K] : rpos[K - 1];
        sh.bcast[0] = cut;
    }
    __syncthreads();
    const int32_t cut = sh.bcast[0];
    __syncthreads();
    return cut;
}

__global__ void k_iota(int32_t* __restrict__ perm, int64_t n) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) perm[g] = (int32_t)g;
}

// One node per workgroup (BVH.hpp:136-189): box, then for an internal node the
// centroid bounds, split axis and the nth_element order of its range.  A range
// longer than lds_cap runs its partition rounds on device memory with the
// whole workgroup (partition_round) until the live range fits in LDS, where one
// lane finishes the select.
__global__ void __launch_bounds__(1024) k_bvh_level(const BvhTask* __restrict__ tasks, const double* __restrict__ box,
                                                   const double* __restrict__ cent, int32_t* __restrict__ perm,
                                                   KeyId* __restrict__ work, int32_t* __restrict__ scr_l,
                                                   int32_t* __restrict__ scr_r, BvhNode* __restrict__ nodes,
                                                   double* __restrict__ leaf_box, int lds_cap) {
    extern __shared__ KeyId lds[];
    __shared__ BlockShared sh;
    const BvhTask T = tasks[blockIdx.x];
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;

    double lo[3], hi[3];
    fold_bounds(box, 6, 3, perm, T.begin, T.n, sh, lo, hi);   // BVH.hpp:141-143
    if (T.right < 0) {                                        // leaf, BVH.hpp:145-152
        if (tid == 0) {
            BvhNode nd;
            for (int a = 0; a < 3; ++a) { nd.lo[a] = lo[a]; nd.hi[a] = hi[a]; }
            nd.right = -1; nd.start = T.begin; nd.count = T.n; nd.axis = 0;
            nodes[T.node] = nd;
            if (leaf_box)
                for (int a = 0; a < 3; ++a) { leaf_box[6 * (size_t)T.begin + a] = lo[a]; leaf_box[6 * (size_t)T.begin + 3 + a] = hi[a]; }
        }
        return;
    }
    double cmin[3], cmax[3];
    fold_bounds(cent, 3, 0, perm, T.begin, T.n, sh, cmin, cmax);   // BVH.hpp:155-161
    const double spread[3] = { cmax[0] - cmin[0], cmax[1] - cmin[1], cmax[2] - cmin[2] };
    int axis = 0;                                                  // std::max_element, BVH.hpp:162-164
    if (spread[axis] < spread[1]) axis = 1;
    if (spread[axis] < spread[2]) axis = 2;

    const int32_t b0 = T.begin, e0 = b0 + T.n, nth = b0 + T.n / 2;   // BVH.hpp:166-175
    if (T.n <= lds_cap) {
        for (int32_t j = tid; j < T.n; j += nt) {
            const int32_t id = perm[b0 + j];
            lds[j] = KeyId{ cent[3 * (size_t)id + axis], id, 0 };
        }
        __syncthreads();
        if (tid == 0) bbrt::sel::nth_element(lds, (int64_t)0, (int64_t)(nth - b0), (int64_t)T.n, KeyLess());
        __syncthreads();
        for (int32_t j = tid; j < T.n; j += nt) perm[b0 + j] = lds[j].id;
    } else {
        for (int32_t j = tid; j < T.n; j += nt) {
            const int32_t id = perm[b0 + j];
            work[b0 + j] = KeyId{ cent[3 * (size_t)id + axis], id, 0 };
        }
        __syncthreads();
        // introselect's rounds (bvh_select.h select_run) while the live range is long
        SelectState st = bbrt::sel::select_begin(b0, nth, e0);
        bool heap_done = false;
        while (st.hi - st.lo >= lds_cap) {
            if (st.depth == 0) {   // the heap_select fallback, by one lane
                if (tid == 0) bbrt::sel::select_run(work, st, (int64_t)0, KeyLess());
                heap_done = true;
                break;
            }
            --st.depth;
            if (tid == 0)
                bbrt::sel::median_to(work, st.lo, st.lo + 1, st.lo + (st.hi - st.lo) / 2, st.hi - 1, KeyLess());
            __syncthreads();
            const int32_t cut = partition_round(work, (int32_t)st.lo, (int32_t)st.hi, scr_l + b0, scr_r + b0, sh);
            if (cut <= st.nth) st.lo = cut;
            else st.hi = cut;
        }
        __syncthreads();
        if (!heap_done) {   // the live range fits in LDS: one lane finishes the select there
            const int64_t m = st.hi - st.lo;
            for (int64_t j = tid; j < m; j += nt) lds[j] = work[st.lo + j];
            __syncthreads();
            if (tid == 0) {
                SelectState s2{ 0, m, st.nth - st.lo, st.depth };
                bbrt::sel::select_run(lds, s2, (int64_t)0, KeyLess());
            }
            __syncthreads();
            for (int64_t j = tid; j < m; j += nt) work[st.lo + j] = lds[j];
            __syncthreads();
        }
        for (int32_t j = tid; j < T.n; j += nt) perm[b0 + j] = work[b0 + j].id;
    }
    if (tid == 0) {   // BVH.hpp:183-187
        BvhNode nd;
        for (int a = 0; a < 3; ++a) { nd.lo[a] = lo[a]; nd.hi[a] = hi[a]; }
        nd.right = T.right; nd.start = 0; nd.count = 0; nd.axis = axis;
        nodes[T.node] = nd;
    }
}

// 4-wide traversal nodes over the finished binary trees (what build_bvh4 does
// on the host): the root and every internal node at even depth hold their
// grandchildren (a leaf child takes one slot) with FP32 boxes widened by the
// part's margin and rounded outward; every other node is all zero.
__global__ void k_bvh4(const BvhTask* __restrict__ tasks, int64_t ntask, const BvhNode* __restrict__ nodes,
                       const double* __restrict__ widen, bbrt::Bvh4Node* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntask) return;
    const BvhTask T = tasks[t];
    bbrt::Bvh4Node q{};
    if (T.depth % 2 == 0 && (T.right >= 0 || T.depth == 0)) {
        const double w = widen[T.part];
        for (int s = 0; s < 4; ++s) {
            q.ref[s] = -1;
            for (int a = 0; a < 3; ++a) { q.lo[a][s] = INFINITY; q.hi[a][s] = -INFINITY; }
        }
        auto put = [&](int s, int32_t k) {
            const BvhNode nd = nodes[k];
            for (int a = 0; a < 3; ++a) { q.lo[a][s] = round_down_f(nd.lo[a] - w); q.hi[a][s] = round_up_f(nd.hi[a] + w); }
            q.ref[s] = nd.right >= 0 ? k : -2 - ((nd.start << 3) | nd.count);
        };
        const BvhNode B = nodes[T.node];
        if (B.right < 0) {
            put(0, T.node);
        } else {
            const int32_t kids[2] = { T.node + 1, B.right };
            q.axes = B.axis;
            for (int h = 0; h < 2; ++h) {
                const BvhNode X = nodes[kids[h]];
                if (X.right < 0) {
                    put(2 * h, kids[h]);
                } else {
                    q.axes |= X.axis << (2 + 2 * h);
                    put(2 * h, kids[h] + 1);
                    put(2 * h + 1, X.right);
                }
            }
        }
    }
    out[T.node] = q;
}

// Nodes of the subtree over m objects (the tree's shape depends on m alone).
int64_t subtree_nodes(int64_t m, int leaf_threshold, std::unordered_map<int64_t, int64_t>& memo) {
    if (m <= leaf_threshold) return 1;
    const auto it = memo.find(m);
    if (it != memo.end()) return it->second;
    const int64_t c = 1 + subtree_nodes(m / 2, leaf_threshold, memo) + subtree_nodes(m - m / 2, leaf_threshold, memo);
    memo.emplace(m, c);
    return c;
}

template <class T>
T* dev_alloc(size_t count, std::vector<void*>& allocs, int64_t& bytes) {
    T* p = nullptr;
    const size_t sz = std::max<size_t>(1, count) * sizeof(T);
    check(hipMalloc(&p, sz), "hipMalloc(bvh)");
    allocs.push_back(p);
    bytes += (int64_t)sz;
    return p;
}

}  // namespace

namespace bbrt {

void build_bvh_device(const double* d_box, const double* d_cent, const std::vector<int64_t>& begin, int leaf_threshold,
                      const double* widen, hipStream_t st, std::vector<void*>& allocs, int64_t& bytes, DeviceTree& out) {
    if (leaf_threshold < 1) throw std::runtime_error("BVH build: leaf threshold below 1");
    if (begin.size() < 2 || begin[0] != 0) throw std::runtime_error("BVH build: bad part offsets");
    const size_t parts = begin.size() - 1;
    const int64_t n_obj = begin.back();
    if (n_obj >= ((int64_t)1 << 30)) throw std::runtime_error("BVH build: too many objects");
    // the nodes, depth by depth
    std::unordered_map<int64_t, int64_t> memo;
    std::vector<std::vector<BvhTask>> by_depth;
    out.roots.assign(parts, 0);
    int64_t next_node = 0;
    for (size_t p = 0; p < parts; ++p) {
        const int64_t n = begin[p + 1] - begin[p];
        if (n <= 0) throw std::runtime_error("BVH build: empty part " + std::to_string(p));
        out.roots[p] = (int32_t)next_node;
        std::vector<BvhTask> todo{ BvhTask{ (int32_t)begin[p], (int32_t)n, (int32_t)next_node, -1, 0, (int32_t)p } };
        next_node += subtree_nodes(n, leaf_threshold, memo);
        if (next_node >= ((int64_t)1 << 30)) throw std::runtime_error("BVH build: too many nodes");
        while (!todo.empty()) {
            BvhTask t = todo.back();
            todo.pop_back();
            if (t.n > leaf_threshold) {
                const int32_t nl = t.n / 2;
                t.right = t.node + 1 + (int32_t)subtree_nodes(nl, leaf_threshold, memo);
                todo.push_back(BvhTask{ t.begin + nl, t.n - nl, t.right, -1, t.depth + 1, t.part });
                todo.push_back(BvhTask{ t.begin, nl, t.node + 1, -1, t.depth + 1, t.part });
            }
            if ((size_t)t.depth >= by_depth.size()) by_depth.resize((size_t)t.depth + 1);
            by_depth[(size_t)t.depth].push_back(t);
        }
    }
    out.n_nodes = next_node;
    std::vector<BvhTask> tasks;
    tasks.reserve((size_t)next_node);
    std::vector<int64_t> level_off{ 0 };
    std::vector<int32_t> level_max;
    for (const auto& lv : by_depth) {
        int32_t mx = 0;
        for (const BvhTask& t : lv) if (t.right >= 0) mx = std::max(mx, t.n);
        tasks.insert(tasks.end(), lv.begin(), lv.end());
        level_off.push_back((int64_t)tasks.size());
        level_max.push_back(mx);
    }
    const bool big = std::any_of(level_max.begin(), level_max.end(), [](int32_t m) { return m > kLdsCap; });

    std::vector<void*> tmp;
    int64_t tmp_bytes = 0;
    try {
        BvhTask* d_tasks = dev_alloc<BvhTask>(tasks.size(), tmp, tmp_bytes);
        KeyId* d_work = big ? dev_alloc<KeyId>((size_t)n_obj, tmp, tmp_bytes) : nullptr;
        int32_t* d_scr_l = big ? dev_alloc<int32_t>((size_t)n_obj, tmp, tmp_bytes) : nullptr;
        int32_t* d_scr_r = big ? dev_alloc<int32_t>((size_t)n_obj, tmp, tmp_bytes) : nullptr;
        check(hipMemcpyAsync(d_tasks, tasks.data(), tasks.size() * sizeof(BvhTask), hipMemcpyHostToDevice, st), "H2D");
        out.bvh = dev_alloc<BvhNode>((size_t)out.n_nodes, allocs, bytes);
        out.leaf = dev_alloc<int32_t>((size_t)n_obj, allocs, bytes);
        out.leaf_box = dev_alloc<double>((size_t)n_obj * 6, allocs, bytes);
        // only leaf starts are ever read; the rest stays zero, as in the host build
        check(hipMemsetAsync(out.leaf_box, 0, (size_t)n_obj * 6 * sizeof(double), st), "hipMemsetAsync");
        hipLaunchKernelGGL(k_iota, dim3((unsigned)((n_obj + 255) / 256)), dim3(256), 0, st, out.leaf, n_obj);
        check(hipGetLastError(), "k_iota");
        for (size_t d = 0; d + 1 < level_off.size(); ++d) {
            const int64_t cnt = level_off[d + 1] - level_off[d];
            const int cap = std::min(kLdsCap, std::max(4, level_max[d]));
            // long ranges run their partition rounds with 256 threads, the rest with one wave
            const int threads = level_max[d] > kLdsCap ? 256 : 64;
            hipLaunchKernelGGL(k_bvh_level, dim3((unsigned)cnt), dim3(threads), (size_t)cap * sizeof(KeyId), st,
                               d_tasks + level_off[d], d_box, d_cent, out.leaf, d_work, d_scr_l, d_scr_r, out.bvh,
                               out.leaf_box, cap);
            check(hipGetLastError(), "k_bvh_level");
        }
        if (widen) {
            double* d_widen = dev_alloc<double>(parts, tmp, tmp_bytes);
            check(hipMemcpyAsync(d_widen, widen, parts * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
            out.bvh4 = dev_alloc<Bvh4Node>((size_t)out.n_nodes, allocs, bytes);
            hipLaunchKernelGGL(k_bvh4, dim3((unsigned)((tasks.size() + 255) / 256)), dim3(256), 0, st, d_tasks,
                               (int64_t)tasks.size(), out.bvh, d_widen, out.bvh4);
            check(hipGetLastError(), "k_bvh4");
        }
        check(hipStreamSynchronize(st), "hipStreamSynchronize(bvh)");
    } catch (...) {
        (void)hipStreamSynchronize(st);
        for (void* p : tmp) (void)hipFree(p);
        throw;
    }
    for (void* p : tmp) (void)hipFree(p);
}

}  // namespace bbrt

extern "C" {

int64_t bbrt_build_bvh_gpu(const double* box, const double* centroid, int64_t n, int32_t leaf_threshold,
                           double* node_box, int32_t* right, int32_t* start, int32_t* count, uint8_t* axis,
                           int32_t* leaf_objects) {
    if (!box || !centroid || n <= 0 || leaf_threshold < 1) {
        g_last_error = "bbrt_build_bvh_gpu: bad arguments";
        return -1;
    }
    std::vector<void*> allocs;
    hipStream_t st = nullptr;
    try {
        check(hipStreamCreate(&st), "hipStreamCreate");
        int64_t bytes = 0;
        double* d_box = dev_alloc<double>((size_t)n * 6, allocs, bytes);
        double* d_cent = dev_alloc<double>((size_t)n * 3, allocs, bytes);
        check(hipMemcpyAsync(d_box, box, (size_t)n * 6 * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
        check(hipMemcpyAsync(d_cent, centroid, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
        bbrt::DeviceTree T;
        bbrt::build_bvh_device(d_box, d_cent, { 0, n }, leaf_threshold, nullptr, st, allocs, bytes, T);
        std::vector<BvhNode> nodes((size_t)T.n_nodes);
        check(hipMemcpy(nodes.data(), T.bvh, nodes.size() * sizeof(BvhNode), hipMemcpyDeviceToHost), "D2H");
        if (leaf_objects) check(hipMemcpy(leaf_objects, T.leaf, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost), "D2H");
        for (size_t k = 0; k < nodes.size(); ++k) {
            const BvhNode& nd = nodes[k];
            if (node_box) for (int a = 0; a < 3; ++a) { node_box[6 * k + a] = nd.lo[a]; node_box[6 * k + 3 + a] = nd.hi[a]; }
            if (right) right[k] = nd.right;
            if (start) start[k] = nd.start;
            if (count) count[k] = nd.count;
            if (axis) axis[k] = (uint8_t)nd.axis;
        }
        for (void* p : allocs) (void)hipFree(p);
        (void)hipStreamDestroy(st);
        return T.n_nodes;
    } catch (const std::exception& e) {
        for (void* p : allocs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
        fail(e);
        return -1;
    }
}

}  // extern "C"
