// Triplet ranking order of determine_three_noncollinear_vertices
// (math.cpp:289-389): the reference collects every (i < j < k) vertex triplet
// of a projected polygon whose orient2d is not zero, in lexicographic order,
// sorts them with std::sort and the comparator "greater |orient2d| first"
// (math.cpp:373-381), and takes the first.  Equal keys are common (the four
// triangles of a square), and which of them comes first depends on std::sort's
// element moves, so the sort is restated step for step as GCC's libstdc++
// performs it:
//   - introsort loop: while a range holds more than 16 elements, move the
//     median of (first + 1, middle, last - 1) to the front, partition the rest
//     around it without bound checks, recurse on the right part and loop on
//     the left; a range whose depth budget (2 floor(log2 n) for the whole
//     array) runs out is heap sorted (std::__partial_sort over the range);
//   - final insertion sort: guarded over the first 16 elements, unguarded
//     after them (or guarded over the whole array when it holds <= 16).
// The functions are usable on the host too, where a test holds them to
// libstdc++'s own std::sort on adversarial inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcgpu {

struct Trip {
    double a;     // |orient2d|
    uint64_t ijk; // the triplet's vertex indices, 16 bits each (i | j << 16 | k << 32)
};

__host__ __device__ inline bool trip_less(const Trip& x, const Trip& y) { return x.a > y.a; }

__host__ __device__ inline void trip_swap(Trip* x, Trip* y)
{
    Trip t = *x;
    *x = *y;
    *y = t;
}

__host__ __device__ inline void trip_linear_insert(Trip* last)
{
    Trip val = *last;
    Trip* next = last - 1;
    while (trip_less(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

__host__ __device__ inline void trip_insertion_sort(Trip* first, Trip* last)
{
    if (first == last) return;
    for (Trip* i = first + 1; i != last; ++i) {
        if (trip_less(*i, *first)) {
            Trip val = *i;
            for (Trip* p = i; p != first; --p) *p = *(p - 1);
            *first = val;
        } else {
            trip_linear_insert(i);
        }
    }
}

// binary max-heap under trip_less: the hole at `hole` moves down along the
// larger child to a leaf, then `value` rises from there no higher than `top`
__host__ __device__ inline void trip_adjust_heap(Trip* first, long hole, long len, Trip value)
{
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (trip_less(first[child], first[child - 1])) child--;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    long parent = (hole - 1) / 2;
    while (hole > top && trip_less(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

// std::__partial_sort(first, last, last): build the heap, then pop it
__host__ __device__ inline void trip_heap_sort(Trip* first, Trip* last)
{
    const long len = last - first;
    if (len >= 2)
        for (long parent = (len - 2) / 2;; --parent) {
            trip_adjust_heap(first, parent, len, first[parent]);
            if (parent == 0) break;
        }
    while (last - first > 1) {
        --last;
        const Trip value = *last;
        *last = *first;
        trip_adjust_heap(first, 0, last - first, value);
    }
}

// std::__move_median_to_first(first, first + 1, mid, last - 1)
__host__ __device__ inline void trip_median_to_first(Trip* f, Trip* l)
{
    Trip* mid = f + (l - f) / 2;
    Trip *a = f + 1, *b = mid, *c = l - 1;
    if (trip_less(*a, *b)) {
        if (trip_less(*b, *c)) trip_swap(f, b);
        else if (trip_less(*a, *c)) trip_swap(f, c);
        else trip_swap(f, a);
    } else if (trip_less(*a, *c)) {
        trip_swap(f, a);
    } else if (trip_less(*b, *c)) {
        trip_swap(f, c);
    } else {
        trip_swap(f, b);
    }
}

// std::__unguarded_partition(f + 1, l, f): the median-of-three move leaves an
// element on each side that stops each scan, so neither runs off the range
__host__ __device__ inline Trip* trip_unguarded_partition(Trip* f, Trip* l)
{
    Trip* lo = f + 1;
    Trip* hi = l;
    for (;;) {
        while (trip_less(*lo, *f)) ++lo;
        --hi;
        while (trip_less(*f, *hi)) --hi;
        if (!(lo < hi)) return lo;
        trip_swap(lo, hi);
        ++lo;
    }
}

__host__ __device__ inline int trip_depth_budget(long n)
{
    int lg = 0;
    while ((1L << (lg + 1)) <= n) ++lg;
    return 2 * lg;
}

// Sorts [first, last) as libstdc++'s std::sort would.  The introsort
// recursion is kept on an explicit stack: ranges are independent, so the
// order in which they are finished does not change the result.  Returns the
// number of ranges that fell back to the heap sort.
__host__ __device__ inline int trip_sort(Trip* first, Trip* last)
{
    const long n = last - first;
    if (n <= 0) return 0;
    // a range is split only while longer than 16 and its depth budget lasts,
    // and the budget is at most 2 log2(n) < 128 levels
    Trip* seg_first[130];
    Trip* seg_last[130];
    int seg_depth[130];
    int top = 0, heaps = 0;
    seg_first[top] = first;
    seg_last[top] = last;
    seg_depth[top] = trip_depth_budget(n);
    ++top;
    while (top > 0) {
        --top;
        Trip* f = seg_first[top];
        Trip* l = seg_last[top];
        int depth = seg_depth[top];
        while (l - f > 16) {
            if (depth == 0) {
                trip_heap_sort(f, l);
                ++heaps;
                break;
            }
            --depth;
            trip_median_to_first(f, l);
            Trip* lo = trip_unguarded_partition(f, l);
            // the right part [lo, l) first, the left part [f, lo) later; every
            // pushed range has a smaller depth budget than its parent, so the
            // stack holds at most one range per budget level
            seg_first[top] = f;
            seg_last[top] = lo;
            seg_depth[top] = depth;
            ++top;
            f = lo;
        }
    }
    if (n > 16) {
        trip_insertion_sort(first, first + 16);
        for (Trip* i = first + 16; i != last; ++i) trip_linear_insert(i);
    } else {
        trip_insertion_sort(first, last);
    }
    return heaps;
}

// The element std::sort puts first, without the rest of the sort.  The
// introsort recurses on the right part of each partition and loops on the
// left one, and a right part never moves an element of the left, so only the
// chain of left parts [first, cut) decides position 0.  That chain ends in a
// range of at most 16 elements or in a heap sort of the range (depth budget
// spent).  Every element right of a cut is not greater than every element
// left of it, and the final insertion sort moves an element only past
// strictly smaller ones, so position 0 then holds the first element of the
// largest key within the chain's last range (or the heap sort's first).
// Expected work is linear in n instead of n log n.  [first, last) is
// permuted on the way.
__host__ __device__ inline Trip trip_first(Trip* first, Trip* last)
{
    Trip* l = last;
    int depth = trip_depth_budget(last - first);
    while (l - first > 16) {
        if (depth == 0) {
            trip_heap_sort(first, l);
            return *first;
        }
        --depth;
        trip_median_to_first(first, l);
        l = trip_unguarded_partition(first, l);
    }
    Trip best = *first;
    for (Trip* p = first + 1; p < l; ++p)
        if (trip_less(*p, best)) best = *p;
    return best;
}

} // namespace mcgpu
