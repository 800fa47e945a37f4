// bvh_select.h -- the element order std::nth_element leaves, reproduced for the
// device BVH build.
//
// BVH::build (include/BVH.hpp:173-179) splits each node's objects with
// std::nth_element at the middle, and the order it leaves the objects in is
// the leaf order of the reference's tree (leaf_objects_, :151-156) and the input
// order of every node below.  The reference is built with GCC's libstdc++,
// whose nth_element is introselect; this file states that algorithm (as
// published in libstdc++'s bits/stl_algo.h and bits/stl_heap.h) step for step,
// over an array of elements with a strict-weak-order `less`:
//   nth_element:    depth limit 2 * floor(log2(n)), then introselect;
//   introselect:    while more than 3 elements remain: at the depth limit,
//                   heap_select over [first, nth + 1) and swap first with nth;
//                   else the median of (first + 1, mid, last - 1) moves to first
//                   and the unguarded partition around it picks the side
//                   holding nth; then insertion sort of the last <= 3;
//   heap_select:    make_heap on [first, middle), then every later element
//                   smaller than the top replaces it (pop_heap);
//   heaps:          the sift-down-to-a-leaf-then-sift-up adjust_heap.
// Every comparison and move is the library's, so the result is the library's
// permutation, not merely a valid partition (tests/test_nth_element.py checks
// it against std::nth_element on random, duplicate-heavy and adversarial
// inputs that reach the heap_select fallback).
//
// The partition scans are bounded by the range: with a strict weak order (any
// finite keys) the median-of-three sentinels stop them first, so the bounds
// change nothing; they keep an inconsistent order (NaN keys) in the range.
//
// Compiled for the device (bbrt_bvh.hip) and, for the test, on the host with
// __host__, __device__ and __forceinline__ defined away.
#pragma once
#include <stdint.h>

namespace bbrt {
namespace sel {

// An element of the device build: the object's centroid coordinate on the
// node's split axis, and the object's index.
struct KeyId {
    double key;
    int32_t id;
    int32_t pad;
};
struct KeyLess {
    __host__ __device__ __forceinline__ bool operator()(const KeyId& a, const KeyId& b) const { return a.key < b.key; }
};

template <class E>
__host__ __device__ __forceinline__ void swap_at(E* a, int64_t i, int64_t j) {
    const E t = a[i];
    a[i] = a[j];
    a[j] = t;
}

__host__ __device__ __forceinline__ int floor_log2(int64_t n) {
    int r = -1;
    while (n > 0) { n >>= 1; ++r; }
    return r;
}

// The median of a[x], a[y], a[z] is swapped into a[res].
template <class E, class Less>
__host__ __device__ inline void median_to(E* a, int64_t res, int64_t x, int64_t y, int64_t z, Less lt) {
    int64_t m;
    if (lt(a[x], a[y])) m = lt(a[y], a[z]) ? y : (lt(a[x], a[z]) ? z : x);
    else m = lt(a[x], a[z]) ? x : (lt(a[y], a[z]) ? z : y);
    swap_at(a, res, m);
}

// Partition of [lo, hi) around the pivot a[pv] (pv < lo), scanning inward and
// swapping the out-of-place pairs; returns where the left scan stopped last.
template <class E, class Less>
__host__ __device__ inline int64_t partition_around(E* a, int64_t lo, int64_t hi, int64_t pv, Less lt) {
    const int64_t top = hi - 1;
    for (;;) {
        while (lo < top && lt(a[lo], a[pv])) ++lo;
        --hi;
        while (hi > pv && lt(a[pv], a[hi])) --hi;
        if (!(lo < hi)) return lo;
        swap_at(a, lo, hi);
        ++lo;
    }
}

// The same partition round from ranks, the form the device runs in parallel.
// Let L be the positions in [lo, hi) whose element is not below the pivot
// (the left scan's stoppers), ascending, and R the positions in [pv, hi) whose
// element is not above it (the right scan's stoppers, the pivot itself
// included), descending.  Between swaps each scan sees original elements
// until it reaches a place the other scan has filled, which holds a stopper
// for it; so the k-th swap exchanges L[k] and R[k], swaps happen exactly for
// the prefix of k with L[k] < R[k] (L rises, R falls), say k < K, and the left
// scan then stops at min(L[K], R[K-1]) -- at L[0] when K = 0.  lpos and rpos
// receive L and R (room for hi - pv entries each).
template <class E, class Less>
__host__ __device__ inline int64_t partition_by_ranks(E* a, int64_t lo, int64_t hi, int64_t pv, Less lt,
                                                      int64_t* lpos, int64_t* rpos) {
    int64_t nl = 0, nr = 0;
    for (int64_t p = lo; p < hi; ++p)
        if (!lt(a[p], a[pv])) lpos[nl++] = p;
    for (int64_t p = hi - 1; p >= pv; --p)
        if (!lt(a[pv], a[p])) rpos[nr++] = p;
    int64_t K = 0;
    while (K < nl && K < nr && lpos[K] < rpos[K]) ++K;
    for (int64_t k = 0; k < K; ++k) swap_at(a, lpos[k], rpos[k]);
    if (K == 0) return nl > 0 ? lpos[0] : hi - 1;
    return (K < nl && lpos[K] < rpos[K - 1]) ? lpos[K] : rpos[K - 1];
}

template <class E, class Less>
__host__ __device__ inline void insertion_sort(E* a, int64_t lo, int64_t hi, Less lt) {
    if (lo == hi) return;
    for (int64_t i = lo + 1; i != hi; ++i) {
        const E v = a[i];
        if (lt(v, a[lo])) {
            for (int64_t j = i; j > lo; --j) a[j] = a[j - 1];
            a[lo] = v;
        } else {
            int64_t j = i;
            while (lt(v, a[j - 1])) { a[j] = a[j - 1]; --j; }
            a[j] = v;
        }
    }
}

// Heaps over a[base .. base + len), positions relative to base.
template <class E, class Less>
__host__ __device__ inline void heap_sift_up(E* a, int64_t base, int64_t hole, int64_t top, E v, Less lt) {
    int64_t parent = (hole - 1) / 2;
    while (hole > top && lt(a[base + parent], v)) {
        a[base + hole] = a[base + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[base + hole] = v;
}

template <class E, class Less>
__host__ __device__ inline void heap_adjust(E* a, int64_t base, int64_t hole, int64_t len, E v, Less lt) {
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (lt(a[base + child], a[base + child - 1])) --child;
        a[base + hole] = a[base + child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[base + hole] = a[base + child - 1];
        hole = child - 1;
    }
    heap_sift_up(a, base, hole, top, v, lt);
}

template <class E, class Less>
__host__ __device__ inline void heap_select(E* a, int64_t lo, int64_t mid, int64_t hi, Less lt) {
    const int64_t len = mid - lo;
    if (len >= 2) {
        for (int64_t parent = (len - 2) / 2;; --parent) {
            heap_adjust(a, lo, parent, len, a[lo + parent], lt);
            if (parent == 0) break;
        }
    }
    for (int64_t i = mid; i < hi; ++i)
        if (lt(a[i], a[lo])) {
            const E v = a[i];
            a[i] = a[lo];
            heap_adjust(a, lo, (int64_t)0, len, v, lt);
        }
}

// Introselect state: the live range [lo, hi) that holds nth, and the rounds
// left before the heap_select fallback.
struct SelectState {
    int64_t lo, hi, nth;
    int depth;
};

__host__ __device__ __forceinline__ SelectState select_begin(int64_t lo, int64_t nth, int64_t hi) {
    return SelectState{ lo, hi, nth, 2 * floor_log2(hi - lo) };
}

// Runs introselect on `st` until it finishes (returns true) or, with
// stop_below > 0, until the live range is shorter than stop_below elements
// (returns false; the state resumes on a copy of the range).
template <class E, class Less>
__host__ __device__ inline bool select_run(E* a, SelectState& st, int64_t stop_below, Less lt) {
    if (st.lo == st.hi || st.nth == st.hi) return true;
    while (st.hi - st.lo > 3) {
        if (stop_below > 0 && st.hi - st.lo < stop_below) return false;
        if (st.depth == 0) {
            heap_select(a, st.lo, st.nth + 1, st.hi, lt);
            swap_at(a, st.lo, st.nth);
            return true;
        }
        --st.depth;
        median_to(a, st.lo, st.lo + 1, st.lo + (st.hi - st.lo) / 2, st.hi - 1, lt);
        const int64_t cut = partition_around(a, st.lo + 1, st.hi, st.lo, lt);
        if (cut <= st.nth) st.lo = cut;
        else st.hi = cut;
    }
    insertion_sort(a, st.lo, st.hi, lt);
    return true;
}

template <class E, class Less>
__host__ __device__ inline void nth_element(E* a, int64_t lo, int64_t nth, int64_t hi, Less lt) {
    SelectState st = select_begin(lo, nth, hi);
    select_run(a, st, 0, lt);
}

}  // namespace sel
}  // namespace bbrt
