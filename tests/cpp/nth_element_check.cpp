// Checks bbrt::sel::nth_element (billardballraytracing3d_amd/csrc/bvh_select.h,
// the device BVH build's split) against std::nth_element: the whole permutation
// must be the same, element for element, since it becomes the BVH's leaf order.
// The rank form of a partition round (partition_by_ranks, what the device runs
// in parallel on long ranges) must give the scans' array and cut every round.
// Inputs: random keys (distinct, few distinct values, many equal), sorted,
// reversed, organ-pipe, and inputs built by McIlroy's adversary ("A killer
// adversary for quicksort", 1999) against the median-of-three partition, which
// drive introselect into its heap_select fallback.  Also the BVH build's own
// use: median splits applied recursively, as BVH::build does.
//
// Built by tests/test_nth_element.py with g++ and __host__, __device__,
// __forceinline__ defined away; prints one summary line, exits 1 on a mismatch.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <vector>

#include "bvh_select.h"

using bbrt::sel::KeyId;
using bbrt::sel::KeyLess;

static int g_fail = 0;
static long g_cases = 0;

static std::vector<KeyId> with_keys(const std::vector<double>& k) {
    std::vector<KeyId> a(k.size());
    for (size_t i = 0; i < k.size(); ++i) a[i] = KeyId{ k[i], (int32_t)i, 0 };
    return a;
}

static bool same(const std::vector<KeyId>& x, const std::vector<KeyId>& y) {
    for (size_t i = 0; i < x.size(); ++i)
        if (x[i].id != y[i].id) return false;
    return true;
}

// introselect whose partition rounds are computed from ranks (partition_by_ranks), as the device runs the
// rounds of long ranges; everything else is select_run's
static void nth_element_by_ranks(std::vector<KeyId>& v, int64_t nth) {
    KeyId* a = v.data();
    const int64_t n = (int64_t)v.size();
    if (n == 0 || nth == n) return;
    std::vector<int64_t> lpos((size_t)n), rpos((size_t)n);
    bbrt::sel::SelectState st = bbrt::sel::select_begin(0, nth, n);
    while (st.hi - st.lo > 3 && st.depth > 0) {
        --st.depth;
        bbrt::sel::median_to(a, st.lo, st.lo + 1, st.lo + (st.hi - st.lo) / 2, st.hi - 1, KeyLess());
        std::vector<KeyId> w(v);
        const int64_t c0 = bbrt::sel::partition_around(w.data(), st.lo + 1, st.hi, st.lo, KeyLess());
        const int64_t cut = bbrt::sel::partition_by_ranks(a, st.lo + 1, st.hi, st.lo, KeyLess(), lpos.data(), rpos.data());
        if (cut != c0 || !same(v, w)) {
            if (g_fail < 10) std::printf("MISMATCH partition round n=%lld\n", (long long)n);
            ++g_fail;
        }
        if (cut <= st.nth) st.lo = cut;
        else st.hi = cut;
    }
    bbrt::sel::select_run(a, st, 0, KeyLess());
}

static void check(const std::vector<double>& keys, int64_t nth, const char* what) {
    std::vector<KeyId> a = with_keys(keys), b = a, c = a;
    bbrt::sel::nth_element(a.data(), 0, nth, (int64_t)a.size(), KeyLess());
    std::nth_element(b.begin(), b.begin() + nth, b.end(), [](const KeyId& x, const KeyId& y) { return x.key < y.key; });
    nth_element_by_ranks(c, nth);
    ++g_cases;
    if (!same(a, b) || !same(c, b)) {
        if (g_fail < 10) std::printf("MISMATCH %s n=%zu nth=%lld\n", what, keys.size(), (long long)nth);
        ++g_fail;
    }
}

// Recursive median splits as in BVH::build (BVH.hpp:173-185), the left part first.
static void check_tree(const std::vector<double>& keys, const char* what) {
    std::vector<KeyId> a = with_keys(keys), b = a;
    struct R { int64_t lo, hi; };
    std::vector<R> todo{ { 0, (int64_t)keys.size() } };
    while (!todo.empty()) {
        const R r = todo.back();
        todo.pop_back();
        if (r.hi - r.lo <= 4) continue;
        const int64_t mid = r.lo + (r.hi - r.lo) / 2;
        bbrt::sel::nth_element(a.data(), r.lo, mid, r.hi, KeyLess());
        std::nth_element(b.begin() + r.lo, b.begin() + mid, b.begin() + r.hi,
                         [](const KeyId& x, const KeyId& y) { return x.key < y.key; });
        todo.push_back({ mid, r.hi });
        todo.push_back({ r.lo, mid });
    }
    ++g_cases;
    if (!same(a, b)) {
        if (g_fail < 10) std::printf("MISMATCH tree %s n=%zu\n", what, keys.size());
        ++g_fail;
    }
}

// McIlroy's adversary: keys are decided lazily during one run of the algorithm
// so that the pivot candidates come out as small as possible.
struct Adversary {
    std::vector<int> val;
    int gas, nsolid = 0, candidate = 0;
    explicit Adversary(int n) : val((size_t)n, n), gas(n) {}
    bool operator()(int x, int y) {
        if (val[x] == gas && val[y] == gas) val[x == candidate ? x : y] = nsolid++;
        if (val[x] == gas) candidate = x;
        else if (val[y] == gas) candidate = y;
        return val[x] < val[y];
    }
};

// Partition rounds introselect would run without its depth limit (the header's
// own steps, run to the end): more than 2 * floor(log2 n) means the real run
// took the heap_select fallback.
static int free_rounds(const std::vector<double>& keys, int64_t nth) {
    std::vector<KeyId> a = with_keys(keys);
    int64_t lo = 0, hi = (int64_t)a.size();
    int rounds = 0;
    while (hi - lo > 3) {
        ++rounds;
        bbrt::sel::median_to(a.data(), lo, lo + 1, lo + (hi - lo) / 2, hi - 1, KeyLess());
        const int64_t cut = bbrt::sel::partition_around(a.data(), lo + 1, hi, lo, KeyLess());
        if (cut <= nth) lo = cut;
        else hi = cut;
    }
    return rounds;
}

int main() {
    std::mt19937_64 rng(20261204);
    // every size up to 300 at the BVH's nth, several key distributions
    for (int n = 0; n <= 300; ++n) {
        for (int kind = 0; kind < 6; ++kind) {
            std::vector<double> k((size_t)n);
            for (int i = 0; i < n; ++i) {
                switch (kind) {
                    case 0: k[i] = std::uniform_real_distribution<double>(-1, 1)(rng); break;
                    case 1: k[i] = (double)(rng() % 4); break;                       // few distinct
                    case 2: k[i] = 0.5; break;                                       // all equal
                    case 3: k[i] = i; break;                                         // sorted
                    case 4: k[i] = n - i; break;                                     // reversed
                    default: k[i] = (i < n / 2) ? i : n - i; break;                  // organ pipe
                }
            }
            if (n > 0) check(k, n / 2, "kinds");
            if (n > 1) check(k, (int64_t)(rng() % (uint64_t)n), "kinds, any nth");
        }
    }
    // signed zeros compare equal: their order is the library's
    for (int n = 5; n < 200; n += 7) {
        std::vector<double> k((size_t)n);
        for (int i = 0; i < n; ++i) k[i] = (rng() & 1) ? 0.0 : -0.0;
        check(k, n / 2, "signed zeros");
    }
    // larger random arrays and whole median-split trees
    for (int rep = 0; rep < 60; ++rep) {
        const int n = 1000 + (int)(rng() % 20000);
        std::vector<double> k((size_t)n);
        const int distinct = (rep % 3 == 0) ? 7 : 1 << 30;
        for (int i = 0; i < n; ++i) k[i] = (double)(rng() % (uint64_t)distinct) * 0.25;
        check(k, n / 2, "large");
        check_tree(k, "large");
    }
    // adversarial inputs: the heap_select fallback
    int fallback = 0;
    for (int n = 8; n <= 4000; n = n * 5 / 4 + 1) {
        for (int which = 0; which < 2; ++which) {
            const int64_t nth = which ? n / 2 : n - 2;
            Adversary adv(n);
            std::vector<int> ids((size_t)n);
            for (int i = 0; i < n; ++i) ids[i] = i;
            bbrt::sel::nth_element(ids.data(), 0, nth, (int64_t)n, std::ref(adv));
            std::vector<double> k((size_t)n);
            for (int i = 0; i < n; ++i) k[i] = adv.val[i];
            if (free_rounds(k, nth) > 2 * bbrt::sel::floor_log2(n)) ++fallback;
            check(k, nth, "adversary");
            check_tree(k, "adversary");
        }
    }
    std::printf("cases %ld mismatches %d heap_select_inputs %d\n", g_cases, g_fail, fallback);
    return (g_fail == 0 && fallback > 0) ? 0 : 1;
}
