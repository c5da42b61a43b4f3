// std::sort's exact element order on the device.
//
// The reference's top_k_infeasible (branching.h:43-65) orders candidates
// with std::sort on a non-strict key (distance), so candidates of equal
// distance leave in whatever order the standard library's introsort leaves
// them; the host driver keeps that order (it calls std::sort), and a search
// run on the device must reproduce it to expand the same tree.  This is the
// same introsort as the GNU C++ library's std::sort, written for one lane:
// median-of-three pivot moved to the front, unguarded partition, recursion
// on the right part and iteration on the left while more than 16 elements
// remain, heap sort once the depth budget 2 floor(log2 n) runs out, and a
// final insertion sort (guarded over the first 16 elements, unguarded after).
// The recursion becomes an explicit stack of (first, last, depth) ranges.
// tests/cpp/stdsort_test.cpp checks the result against std::sort on the
// host, element for element, on inputs with many equal keys.
#pragma once

#include <hip/hip_runtime.h>

namespace gb {

template <class T>
__host__ __device__ inline void ss_swap(T* a, T* b) {
    T t = *a;
    *a = *b;
    *b = t;
}

template <class T, class Less>
__host__ __device__ inline void ss_move_median_to_first(T* result, T* a, T* b, T* c, Less less) {
    if (less(*a, *b)) {
        if (less(*b, *c)) ss_swap(result, b);
        else if (less(*a, *c)) ss_swap(result, c);
        else ss_swap(result, a);
    } else if (less(*a, *c)) {
        ss_swap(result, a);
    } else if (less(*b, *c)) {
        ss_swap(result, c);
    } else {
        ss_swap(result, b);
    }
}

template <class T, class Less>
__host__ __device__ inline T* ss_unguarded_partition(T* first, T* last, T* pivot, Less less) {
    while (true) {
        while (less(*first, *pivot)) ++first;
        --last;
        while (less(*pivot, *last)) --last;
        if (!(first < last)) return first;
        ss_swap(first, last);
        ++first;
    }
}

// Heap of [first, first + len): sift the hole at `hole` down to a leaf
// (larger child, the right one unless the left is larger), then the value
// back up towards `top`.
template <class T, class Less>
__host__ __device__ inline void ss_adjust_heap(T* first, long hole, long len, T value, Less less) {
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (less(first[child], first[child - 1])) --child;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    long parent = (hole - 1) / 2;
    while (hole > top && less(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

template <class T, class Less>
__host__ __device__ inline void ss_heap_sort(T* first, T* last, Less less) {
    const long len = last - first;
    if (len >= 2) {
        for (long parent = (len - 2) / 2;; --parent) {
            ss_adjust_heap(first, parent, len, first[parent], less);
            if (parent == 0) break;
        }
    }
    while (last - first > 1) {
        --last;
        T value = *last;
        *last = *first;
        ss_adjust_heap(first, 0L, (long)(last - first), value, less);
    }
}

template <class T, class Less>
__host__ __device__ inline void ss_unguarded_linear_insert(T* last, Less less) {
    T value = *last;
    T* next = last - 1;
    while (less(value, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = value;
}

template <class T, class Less>
__host__ __device__ inline void ss_insertion_sort(T* first, T* last, Less less) {
    if (first == last) return;
    for (T* i = first + 1; i != last; ++i) {
        if (less(*i, *first)) {
            T value = *i;
            for (T* p = i; p != first; --p) *p = *(p - 1);
            *first = value;
        } else {
            ss_unguarded_linear_insert(i, less);
        }
    }
}

// Sorts [first, last) into std::sort's order.  Up to 2^20 elements (the
// explicit stack holds one range per level of the depth budget).
template <class T, class Less>
__host__ __device__ inline void std_order_sort(T* first, T* last, Less less) {
    constexpr long kThreshold = 16;
    const long n = last - first;
    if (n <= 1) return;
    int lg = 0;
    while ((2L << lg) <= n) ++lg;  // floor(log2 n)
    struct Range {
        T* first;
        T* last;
        int depth;
    };
    Range stack[48];
    int top = 0;
    stack[top++] = {first, last, 2 * lg};
    while (top > 0) {
        Range r = stack[--top];
        // the loop of one introsort call: partition, recurse on the right
        // part (pushed), continue with the left
        while (r.last - r.first > kThreshold) {
            if (r.depth == 0) {
                ss_heap_sort(r.first, r.last, less);
                break;
            }
            --r.depth;
            T* mid = r.first + (r.last - r.first) / 2;
            ss_move_median_to_first(r.first, r.first + 1, mid, r.last - 1, less);
            T* cut = ss_unguarded_partition(r.first + 1, r.last, r.first, less);
            // The library sorts the right part [cut, last) completely before
            // continuing with the left one; ranges are disjoint, so the order
            // in which they are sorted does not change the result.
            stack[top++] = {cut, r.last, r.depth};
            r.last = cut;
        }
    }
    if (n > kThreshold) {
        ss_insertion_sort(first, first + kThreshold, less);
        for (T* i = first + kThreshold; i != last; ++i) ss_unguarded_linear_insert(i, less);
    } else {
        ss_insertion_sort(first, last, less);
    }
}

}  // namespace gb
