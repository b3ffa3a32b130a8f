// Depth sort of one sample's hit list, element for element the sequence of
// steps GCC 11's libstdc++ std::sort takes (introsort with median-of-three
// pivots, a heap sort below the depth limit, insertion sorts on runs of 16 or
// fewer; bits/stl_algo.h, bits/stl_heap.h).  CqImagePixel::Combine sorts hits
// with std::sort and CqAscendingDepthSort (imagepixel.cpp:129-142, 167), which
// is not stable, so hits of equal depth end up in the order these steps give
// them.  Keys d[] are depths, x[] travels with them.  Used by the sampling
// kernel; plain C++ so that host tests can compare it with std::sort.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define RH_HD __host__ __device__ __forceinline__
#else
#define RH_HD inline
#endif

namespace rhsort {

RH_HD void move_el(float* d, uint32_t* x, int to, int from) {
    d[to] = d[from];
    x[to] = x[from];
}

RH_HD void swap_el(float* d, uint32_t* x, int a, int b) {
    const float td = d[a];
    const uint32_t tx = x[a];
    d[a] = d[b]; x[a] = x[b];
    d[b] = td; x[b] = tx;
}

// __unguarded_linear_insert: an element not less than d[first] sits before.
RH_HD void linear_insert(float* d, uint32_t* x, int last) {
    const float vd = d[last];
    const uint32_t vx = x[last];
    int next = last - 1;
    while (vd < d[next]) {
        move_el(d, x, last, next);
        last = next;
        --next;
    }
    d[last] = vd;
    x[last] = vx;
}

// __insertion_sort on [first, last)
RH_HD void insertion_sort(float* d, uint32_t* x, int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        if (d[i] < d[first]) {
            const float vd = d[i];
            const uint32_t vx = x[i];
            for (int k = i; k > first; --k) move_el(d, x, k, k - 1);
            d[first] = vd;
            x[first] = vx;
        } else {
            linear_insert(d, x, i);
        }
    }
}

// __push_heap on the range starting at first
RH_HD void push_heap(float* d, uint32_t* x, int first, int hole, int top, float vd, uint32_t vx) {
    int parent = (hole - 1) / 2;
    while (hole > top && d[first + parent] < vd) {
        move_el(d, x, first + hole, first + parent);
        hole = parent;
        parent = (hole - 1) / 2;
    }
    d[first + hole] = vd;
    x[first + hole] = vx;
}

// __adjust_heap
RH_HD void adjust_heap(float* d, uint32_t* x, int first, int hole, int len, float vd, uint32_t vx) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (d[first + child] < d[first + child - 1]) child--;
        move_el(d, x, first + hole, first + child);
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        move_el(d, x, first + hole, first + child - 1);
        hole = child - 1;
    }
    push_heap(d, x, first, hole, top, vd, vx);
}

// __partial_sort(first, last, last): __make_heap then __sort_heap
RH_HD void heap_sort(float* d, uint32_t* x, int first, int last) {
    const int len = last - first;
    if (len >= 2) {
        int parent = (len - 2) / 2;
        while (true) {
            adjust_heap(d, x, first, parent, len, d[first + parent], x[first + parent]);
            if (parent == 0) break;
            parent--;
        }
    }
    while (last - first > 1) {
        --last;
        // __pop_heap(first, last, last)
        const float vd = d[last];
        const uint32_t vx = x[last];
        move_el(d, x, last, first);
        adjust_heap(d, x, first, 0, last - first, vd, vx);
    }
}

// __unguarded_partition_pivot
RH_HD int partition_pivot(float* d, uint32_t* x, int first, int last) {
    const int mid = first + (last - first) / 2;
    const int a = first + 1, b = mid, c = last - 1;
    // __move_median_to_first(first, a, b, c)
    if (d[a] < d[b]) {
        if (d[b] < d[c]) swap_el(d, x, first, b);
        else if (d[a] < d[c]) swap_el(d, x, first, c);
        else swap_el(d, x, first, a);
    } else if (d[a] < d[c]) {
        swap_el(d, x, first, a);
    } else if (d[b] < d[c]) {
        swap_el(d, x, first, c);
    } else {
        swap_el(d, x, first, b);
    }
    // __unguarded_partition(first + 1, last, first)
    int lo = first + 1, hi = last;
    while (true) {
        while (d[lo] < d[first]) ++lo;
        --hi;
        while (d[first] < d[hi]) --hi;
        if (!(lo < hi)) return lo;
        swap_el(d, x, lo, hi);
        ++lo;
    }
}

// std::sort(d, d + n) with x carried along; n <= 64.  `stk` holds 24 ints
// of work stack (caller storage, e.g. LDS).
RH_HD void sort_with(float* d, uint32_t* x, int n, int* stk) {
    if (n <= 0) return;
    int lg = 0;
    while ((2 << lg) <= n) ++lg;          // __lg(n)
    // __introsort_loop: the two parts of a partition are disjoint, so they
    // may be finished in any order; each carries the decremented limit.
    // A stack entry packs first | last << 8 | depth limit << 16.
    int sp = 0;
    stk[0] = 0 | (n << 8) | ((2 * lg) << 16);
    sp = 1;
    while (sp > 0) {
        --sp;
        int first = stk[sp] & 0xff, last = (stk[sp] >> 8) & 0xff, depth = stk[sp] >> 16;
        while (last - first > 16) {
            if (depth == 0) {
                heap_sort(d, x, first, last);
                break;
            }
            --depth;
            const int cut = partition_pivot(d, x, first, last);
            stk[sp] = cut | (last << 8) | (depth << 16);
            ++sp;
            last = cut;
        }
    }
    // __final_insertion_sort
    if (n > 16) {
        insertion_sort(d, x, 0, 16);
        for (int i = 16; i != n; ++i) linear_insert(d, x, i);
    } else {
        insertion_sort(d, x, 0, n);
    }
}

// std::sort(d, d + n) with x carried along for any n (a list longer than
// sort_with takes).  `stk` holds 2 * (2 * lg(n) + 2) ints of work stack:
// (first, last | depth limit << 24) per entry.
RH_HD void sort_long(float* d, uint32_t* x, int n, int* stk) {
    if (n <= 0) return;
    int lg = 0;
    while ((2 << lg) <= n) ++lg;          // __lg(n)
    int sp = 0;
    stk[0] = 0;
    stk[1] = n | ((2 * lg) << 24);
    sp = 1;
    while (sp > 0) {
        --sp;
        int first = stk[2 * sp], last = stk[2 * sp + 1] & 0xffffff, depth = stk[2 * sp + 1] >> 24;
        while (last - first > 16) {
            if (depth == 0) {
                heap_sort(d, x, first, last);
                break;
            }
            --depth;
            const int cut = partition_pivot(d, x, first, last);
            stk[2 * sp] = cut;
            stk[2 * sp + 1] = last | (depth << 24);
            ++sp;
            last = cut;
        }
    }
    if (n > 16) {
        insertion_sort(d, x, 0, 16);
        for (int i = 16; i != n; ++i) linear_insert(d, x, i);
    } else {
        insertion_sort(d, x, 0, n);
    }
}

// std::sort(d, d + n) with x carried along; n <= 64.
RH_HD void sort(float* d, uint32_t* x, int n) {
    int stk[24];
    sort_with(d, x, n, stk);
}

}  // namespace rhsort
