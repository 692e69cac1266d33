// Holds the triplet ranking's sort (mcut_amd/csrc/tripsort.hpp, the code the
// GPU runs, compiled here for the host) and the oracle's restatement of it
// (om_trip_sort_keys, oracle/frontend_oracle.c) to libstdc++'s own std::sort
// with the same comparator, on inputs built to exercise every path of the introsort:
// random keys with many ties, sorted, reversed and organ-pipe keys, and
// McIlroy's adversary ("A Killer Adversary for Quicksort", 1999) run against
// std::sort itself, which drives it into its heap-sort fallback.  Whole
// permutations are compared, not only the first element.  Prints
// "cases <n> mismatches <m> heap_fallbacks <h>".
#include "tripsort.hpp"

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

using mcgpu::Trip;

extern "C" void om_trip_sort_keys(const double* keys, long n, int* order);

static bool same_as_std(const std::vector<double>& keys, int* heaps)
{
    std::vector<Trip> a(keys.size()), b;
    for (size_t i = 0; i < keys.size(); ++i) a[i] = Trip{keys[i], (int32_t)i};
    b = a;
    std::sort(a.begin(), a.end(), [](const Trip& x, const Trip& y) { return mcgpu::trip_less(x, y); });
    *heaps += mcgpu::trip_sort(b.data(), b.data() + b.size());
    std::vector<int> o(keys.size() + 1);
    om_trip_sort_keys(keys.data(), (long)keys.size(), o.data());
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].ijk != b[i].ijk || a[i].ijk != o[i]) return false;
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
    std::printf("cases %d mismatches %d heap_fallbacks %d\n", cases, bad, heaps);
    return bad != 0;
}
