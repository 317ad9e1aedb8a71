// A max-heap over 8-byte elements that performs exactly the element moves of
// libstdc++'s std::priority_queue (std::push_heap / std::pop_heap, i.e.
// __push_heap and __adjust_heap of bits/stl_heap.h), so the order in which
// equivalent elements leave it is the same.  The reference's simplification
// (grid_mscomplex.cpp:387-510) pops equal persistences in that order, and it
// is observable in the output.  Differences from the library: elements are
// packed into one uint64 (ordering key in the high half, payload in the low
// half) and the sift-down selects children without branches.
//
// Ordering: element a has lower priority than b when key(a) > key(b)
// (the reference's comparator on persistence); equal keys are equivalent.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace mscx {

class ExactHeap {
 public:
  static uint64_t make(uint32_t key, uint32_t payload) { return ((uint64_t)key << 32) | payload; }
  static uint32_t key(uint64_t e) { return (uint32_t)(e >> 32); }
  static uint32_t payload(uint64_t e) { return (uint32_t)e; }

  void reserve(size_t n) { a_.reserve(n); }
  bool empty() const { return a_.empty(); }
  size_t size() const { return a_.size(); }
  uint64_t top() const { return a_[0]; }

  // priority_queue::push: push_back, then __push_heap(first, n-1, 0, value)
  void push(uint64_t v) {
    a_.push_back(v);
    sift_up((ptrdiff_t)a_.size() - 1, 0, v);
  }

  // priority_queue::pop: pop_heap (when n > 1: value = last, last = first,
  // __adjust_heap(first, 0, n-1, value)), then pop_back
  void pop() {
    const ptrdiff_t n = (ptrdiff_t)a_.size();
    if (n > 1) {
      uint64_t value = a_[n - 1];
      a_[n - 1] = a_[0];
      adjust(0, n - 1, value);
    }
    a_.pop_back();
  }

 private:
  std::vector<uint64_t> a_;

  // comp(a, b): a has lower priority than b
  static bool comp(uint64_t x, uint64_t y) { return key(x) > key(y); }

  void sift_up(ptrdiff_t hole, ptrdiff_t top, uint64_t value) {
    uint64_t *a = a_.data();
    ptrdiff_t parent = (hole - 1) / 2;
    while (hole > top && comp(a[parent], value)) {
      a[hole] = a[parent];
      hole = parent;
      parent = (hole - 1) / 2;
    }
    a[hole] = value;
  }

  void adjust(ptrdiff_t hole, ptrdiff_t len, uint64_t value) {
    uint64_t *a = a_.data();
    const ptrdiff_t top = hole;
    ptrdiff_t second = hole;
    while (second < (len - 1) / 2) {
      second = 2 * (second + 1);
      // branch-free choice of the higher-priority child (the comparison is
      // data dependent and mispredicts about half the time)
      const uint64_t r = a[second], l = a[second - 1];
      const bool left = comp(r, l);
      second -= left;
      a[hole] = left ? l : r;
      hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
      second = 2 * (second + 1);
      a[hole] = a[second - 1];
      hole = second - 1;
    }
    sift_up(hole, top, value);
  }
};

}  // namespace mscx
