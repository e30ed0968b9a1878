// heap_select.hpp -- exact emulation of the top-k selection done by
// std::partial_sort(first, first+k, last, comp) with comp(a,b) = |a| > |b| on
// (index, value) pairs in index order (reference: faust_prox.hpp:53-69).
//
// partial_sort keeps the elements left in [first, first+k) by its heap
// selection step; the final sort of that range only reorders them, so the
// selected SET is decided by the heap phase alone:
//   1. heapify the first k pairs (parents visited from (k-2)/2 down to 0,
//      each sifted down to a leaf along the child that is not "greater",
//      then bubbled up);
//   2. for each later pair x: if |x| > |top|, x replaces the top and is
//      sifted down/bubbled up the same way.
// The heap is a min-heap on |value| in which ties keep their positions, which
// is what makes the surviving tied elements depend on positions only.
// Written for host and device (plain C++ in a header, no HIP types).
#pragma once

#ifndef FH_HD
#if defined(__HIPCC__)
#define FH_HD __host__ __device__
#else
#define FH_HD
#endif
#endif

namespace fh {

// Heap storage: hidx[0..len) indices, habs[0..len) magnitudes.
template <class Real>
FH_HD inline void heap_adjust(int* hidx, Real* habs, int hole, int len, int vidx, Real vabs) {
  const int top = hole;
  int child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    // comp(right, left) == |right| > |left|: then the left child is taken
    if (habs[child] > habs[child - 1]) child--;
    hidx[hole] = hidx[child];
    habs[hole] = habs[child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    hidx[hole] = hidx[child - 1];
    habs[hole] = habs[child - 1];
    hole = child - 1;
  }
  // bubble up while comp(parent, value) == |parent| > |value|
  int parent = (hole - 1) / 2;
  while (hole > top && habs[parent] > vabs) {
    hidx[hole] = hidx[parent];
    habs[hole] = habs[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  hidx[hole] = vidx;
  habs[hole] = vabs;
}

// Select k of the n magnitudes produced by get_abs(i) (i = 0..n-1), 0 < k < n.
// On return hidx[0..k) holds the selected indices (heap order).
template <class Real, class GetAbs>
FH_HD inline void heap_select(int n, int k, GetAbs get_abs, int* hidx, Real* habs) {
  for (int i = 0; i < k; ++i) {
    hidx[i] = i;
    habs[i] = get_abs(i);
  }
  if (k >= 2) {
    int parent = (k - 2) / 2;
    while (true) {
      int vi = hidx[parent];
      Real va = habs[parent];
      heap_adjust(hidx, habs, parent, k, vi, va);
      if (parent == 0) break;
      parent--;
    }
  }
  for (int i = k; i < n; ++i) {
    Real a = get_abs(i);
    if (a > habs[0]) heap_adjust(hidx, habs, 0, k, i, a);
  }
}

}  // namespace fh
