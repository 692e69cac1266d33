// Holds the triplet ranking's sort (mcut_amd/csrc/tripsort.hpp, the code the
// GPU runs, compiled here for the host) and the oracle's restatement of it
// (om_trip_sort_keys, oracle/frontend_oracle.c) to libstdc++'s own std::sort
// with the same comparator, on inputs built to exercise every path of the introsort:
// random keys with many ties, sorted, reversed and organ-pipe keys, and
// McIlroy's adversary ("A Killer Adversary for Quicksort", 1999) run against
// std::sort itself, which drives it into its heap-sort fallback.  Whole
// permutations are compared, not only the first element.  On the same inputs
// trip_first (the element std::sort puts first, from the left-most chain of
// partitions alone) must equal std::sort's first element, and every partition
// of that chain must come out the same when restated as pairs: the k-th
// element from the left not greater than the pivot swapped with the k-th from
// the right not less than it (the pivot ending that list) while the first
// lies left of the second, the cut then being the first failing pair's left
// position, or the last swapped right position if that is smaller (the form
// the GPU computes a wave at a time).  Prints
// "cases <n> mismatches <m> heap_fallbacks <h> first_mismatches <a> pair_mismatches <b>".
#include "tripsort.hpp"

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

using mcgpu::Trip;

extern "C" void om_trip_sort_keys(const double* keys, long n, int* order);

// the unguarded partition of [f + 1, l) around *f, restated as pairs
static long pairs_partition(std::vector<Trip>& v, long f, long l)
{
    const Trip P = v[f];
    std::vector<long> L, R;
    for (long p = f + 1; p < l; ++p)
        if (!mcgpu::trip_less(v[p], P)) L.push_back(p);
    for (long p = l - 1; p >= f; --p)
        if (!mcgpu::trip_less(P, v[p])) R.push_back(p);
    const long np = (long)std::min(L.size(), R.size());
    long K = np;
    for (long k = 0; k < np; ++k)
        if (!(L[k] < R[k])) {
            K = k;
            break;
        }
    for (long k = 0; k < K; ++k) std::swap(v[L[k]], v[R[k]]);
    if (K == 0) return L[0];
    const long lK = K < (long)L.size() ? L[K] : l;
    return std::min(lK, R[K - 1]);
}

// trip_first against std::sort's first element, and every partition of the
// left-most chain against its restatement as pairs
static void check_first(const std::vector<double>& keys, const Trip& want, int* first_bad, int* pair_bad)
{
    std::vector<Trip> v(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) v[i] = Trip{keys[i], (uint64_t)i};
    if (!v.empty()) {
        std::vector<Trip> w = v;
        const Trip t = mcgpu::trip_first(w.data(), w.data() + w.size());
        if (t.ijk != want.ijk) ++*first_bad;
    }
    long l = (long)v.size();
    int depth = mcgpu::trip_depth_budget(l);
    while (l > 16 && depth > 0) {
        --depth;
        mcgpu::trip_median_to_first(v.data(), v.data() + l);
        std::vector<Trip> u = v;
        const long cut = mcgpu::trip_unguarded_partition(v.data(), v.data() + l) - v.data();
        const long cut2 = pairs_partition(u, 0, l);
        bool same = cut == cut2;
        for (long i = 0; i < l && same; ++i) same = u[i].ijk == v[i].ijk;
        if (!same) ++*pair_bad;
        l = cut;
    }
}

static int first_bad = 0, pair_bad = 0;

static bool same_as_std(const std::vector<double>& keys, int* heaps)
{
    std::vector<Trip> a(keys.size()), b;
    for (size_t i = 0; i < keys.size(); ++i) a[i] = Trip{keys[i], (uint64_t)i};
    b = a;
    std::sort(a.begin(), a.end(), [](const Trip& x, const Trip& y) { return mcgpu::trip_less(x, y); });
    *heaps += mcgpu::trip_sort(b.data(), b.data() + b.size());
    std::vector<int> o(keys.size() + 1);
    om_trip_sort_keys(keys.data(), (long)keys.size(), o.data());
    if (!a.empty()) check_first(keys, a[0], &first_bad, &pair_bad);
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].ijk != b[i].ijk || a[i].ijk != (uint64_t)o[i]) return false;
    return true;
}

// McIlroy's adversary: every element starts as "gas"; comparing two gas
// elements freezes one of them (the current pivot candidate) to the next
// solid value, so the sort's pivots keep landing at an extreme.
static std::vector<double> adversary(int n)
{
    const int gas = n;
    std::vector<int> val(n, gas), idx(n);
    int nsolid = 0, candidate = 0;
    for (int i = 0; i < n; ++i) idx[i] = i;
    auto cmp = [&](int x, int y) {
        if (val[x] == gas && val[y] == gas) {
            if (x == candidate) val[x] = nsolid++;
            else val[y] = nsolid++;
        }
        if (val[x] == gas) candidate = x;
        else if (val[y] == gas) candidate = y;
        return val[x] > val[y]; // the ranking's direction: greater first
    };
    std::sort(idx.begin(), idx.end(), cmp);
    std::vector<double> keys(n);
    for (int i = 0; i < n; ++i) keys[i] = (double)val[i];
    return keys;
}

int main()
{
    std::mt19937_64 rng(7);
    int cases = 0, bad = 0, heaps = 0;
    auto run = [&](const std::vector<double>& k) {
        ++cases;
        if (!same_as_std(k, &heaps)) ++bad;
    };
    for (int n : {1, 2, 3, 5, 15, 16, 17, 18, 31, 32, 33, 100, 560, 1140, 4060, 9880}) {
        for (int ties : {1, 2, 3, 8, 1000000}) {
            std::vector<double> k(n);
            for (auto& x : k) x = (double)(rng() % (uint64_t)ties);
            run(k);
        }
        std::vector<double> k(n);
        for (int i = 0; i < n; ++i) k[i] = i;
        run(k);
        std::reverse(k.begin(), k.end());
        run(k);
        for (int i = 0; i < n; ++i) k[i] = std::min(i, n - 1 - i);
        run(k);
        run(adversary(n));
    }
    for (int n = 17; n < 400; n += 7) run(adversary(n));
    std::printf("cases %d mismatches %d heap_fallbacks %d first_mismatches %d pair_mismatches %d\n", cases, bad, heaps,
                first_bad, pair_bad);
    return bad != 0 || first_bad != 0 || pair_bad != 0;
}
