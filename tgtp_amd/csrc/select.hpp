// select.hpp -- the partial selection std::nth_element performs in
// libstdc++ (GCC 11: introselect = median-of-three unguarded partitioning
// with a 2*floor(log2 n) depth limit, heap selection past the limit and
// insertion sort below four elements).
//
// aesc.h:417 calls std::nth_element(q, q + gamma - 1, q + |q|, pvec[a] >
// pvec[b]); which frontier vertices land in the first gamma slots decides
// edge_max when pvec values tie at the gamma-th value, so the device must
// reproduce the exact permutation, not merely some valid selection.  The
// steps below have to match libstdc++'s; the code is written for this
// project (index-based, one comparator, usable on host and device) and is
// checked against std::nth_element on tie-heavy random inputs by the tests.
#pragma once

#include <cstdint>

#ifndef TGTP_HD
#if defined(__HIPCC__)
#define TGTP_HD __host__ __device__
#else
#define TGTP_HD
#endif
#endif

namespace tgtp_select {

// Comparator: "a before b" iff key(a) > key(b), key = pvec[id].
template <typename Id>
struct Greater {
  const float* key;
  TGTP_HD bool operator()(Id a, Id b) const { return key[a] > key[b]; }
};

template <typename Id, typename Cmp>
TGTP_HD inline void swap_at(Id* q, int64_t i, int64_t j) {
  Id t = q[i];
  q[i] = q[j];
  q[j] = t;
}

// sift `value` down from `hole` in heap q[0, len), then up toward `top`
template <typename Id, typename Cmp>
TGTP_HD void heap_adjust(Id* q, int64_t hole, int64_t len, Id value, Cmp cmp) {
  const int64_t top = hole;
  int64_t child = hole;
  while (child < (len - 1) / 2) {
    child = 2 * (child + 1);
    if (cmp(q[child], q[child - 1])) --child;
    q[hole] = q[child];
    hole = child;
  }
  if ((len & 1) == 0 && child == (len - 2) / 2) {
    child = 2 * (child + 1);
    q[hole] = q[child - 1];
    hole = child - 1;
  }
  int64_t parent = (hole - 1) / 2;
  while (hole > top && cmp(q[parent], value)) {
    q[hole] = q[parent];
    hole = parent;
    parent = (hole - 1) / 2;
  }
  q[hole] = value;
}

template <typename Id, typename Cmp>
TGTP_HD void heap_make(Id* q, int64_t len, Cmp cmp) {
  if (len < 2) return;
  int64_t parent = (len - 2) / 2;
  for (;;) {
    heap_adjust<Id, Cmp>(q, parent, len, q[parent], cmp);
    if (parent == 0) return;
    --parent;
  }
}

// heap selection over q[first, last) keeping the `middle - first` best
template <typename Id, typename Cmp>
TGTP_HD void heap_select(Id* q, int64_t first, int64_t middle, int64_t last, Cmp cmp) {
  Id* base = q + first;
  const int64_t len = middle - first;
  heap_make<Id, Cmp>(base, len, cmp);
  for (int64_t i = middle; i < last; ++i) {
    if (cmp(q[i], base[0])) {
      const Id value = q[i];
      q[i] = base[0];
      heap_adjust<Id, Cmp>(base, 0, len, value, cmp);
    }
  }
}

template <typename Id, typename Cmp>
TGTP_HD void insertion_sort(Id* q, int64_t first, int64_t last, Cmp cmp) {
  if (first == last) return;
  for (int64_t i = first + 1; i != last; ++i) {
    const Id val = q[i];
    if (cmp(val, q[first])) {
      for (int64_t k = i; k > first; --k) q[k] = q[k - 1];
      q[first] = val;
    } else {
      int64_t hole = i;
      int64_t next = i - 1;
      while (cmp(val, q[next])) {
        q[hole] = q[next];
        hole = next;
        --next;
      }
      q[hole] = val;
    }
  }
}

template <typename Id, typename Cmp>
TGTP_HD void median_to_first(Id* q, int64_t result, int64_t a, int64_t b, int64_t c, Cmp cmp) {
  if (cmp(q[a], q[b])) {
    if (cmp(q[b], q[c]))
      swap_at<Id, Cmp>(q, result, b);
    else if (cmp(q[a], q[c]))
      swap_at<Id, Cmp>(q, result, c);
    else
      swap_at<Id, Cmp>(q, result, a);
  } else if (cmp(q[a], q[c])) {
    swap_at<Id, Cmp>(q, result, a);
  } else if (cmp(q[b], q[c])) {
    swap_at<Id, Cmp>(q, result, c);
  } else {
    swap_at<Id, Cmp>(q, result, b);
  }
}

// Hoare-style partition of q[first, last) around the pivot at q[pivot],
// relying on sentinels on both sides (no bounds checks in the scans).
template <typename Id, typename Cmp>
TGTP_HD int64_t partition_unguarded(Id* q, int64_t first, int64_t last, int64_t pivot, Cmp cmp) {
  for (;;) {
    while (cmp(q[first], q[pivot])) ++first;
    --last;
    while (cmp(q[pivot], q[last])) --last;
    if (!(first < last)) return first;
    swap_at<Id, Cmp>(q, first, last);
    ++first;
  }
}

TGTP_HD inline int floor_log2(int64_t n) {
  int r = 0;
  while (n > 1) {
    n >>= 1;
    ++r;
  }
  return r;
}

// Rearranges q[0, len) exactly as std::nth_element(q, q + nth, q + len, cmp).
template <typename Id, typename Cmp>
TGTP_HD void nth_element(Id* q, int64_t nth, int64_t len, Cmp cmp) {
  if (len == 0 || nth == len) return;
  int64_t first = 0, last = len;
  int depth = 2 * floor_log2(len);
  while (last - first > 3) {
    if (depth == 0) {
      heap_select<Id, Cmp>(q, first, nth + 1, last, cmp);
      swap_at<Id, Cmp>(q, first, nth);
      return;
    }
    --depth;
    const int64_t mid = first + (last - first) / 2;
    median_to_first<Id, Cmp>(q, first, first + 1, mid, last - 1, cmp);
    const int64_t cut = partition_unguarded<Id, Cmp>(q, first + 1, last, first, cmp);
    if (cut <= nth)
      first = cut;
    else
      last = cut;
  }
  insertion_sort<Id, Cmp>(q, first, last, cmp);
}

}  // namespace tgtp_select
