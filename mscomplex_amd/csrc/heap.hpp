// A max-heap over 8-byte elements that performs exactly the element moves of
// libstdc++'s std::priority_queue (std::push_heap / std::pop_heap, i.e.
// __push_heap and __adjust_heap of bits/stl_heap.h), so the order in which
// equivalent elements leave it is the same.  The reference's simplification
// (grid_mscomplex.cpp:387-510) pops equal persistences in that order, and it
// is observable in the output.  Differences from the library: elements are
// packed into one uint64 (ordering key in the high half, payload in the low
// half), the sift-down selects children without branches, and it descends two
// levels per step with the four grandchildren loaded before the first choice
// is known (the storage keeps 2n + 8 readable slots for those loads), which
// halves the chain of dependent loads.
//
// Ordering: element a has lower priority than b when key(a) > key(b)
// (the reference's comparator on persistence); equal keys are equivalent.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "hugepages.hpp"

namespace mscx {

class ExactHeap {
 public:
  static uint64_t make(uint32_t key, uint32_t payload) { return ((uint64_t)key << 32) | payload; }
  static uint32_t key(uint64_t e) { return (uint32_t)(e >> 32); }
  static uint32_t payload(uint64_t e) { return (uint32_t)e; }

  // room for n elements; the storage is advised onto huge pages before its
  // first touch (a pop's descent touches a random path of it)
  void reserve(size_t n) {
    if (a_.size() >= 2 * n + 8) return;
    std::vector<uint64_t> b;
    b.reserve(2 * n + 8);
    advise_huge(b.data(), b.capacity() * sizeof(uint64_t));
    b.assign(a_.begin(), a_.end());
    b.resize(2 * n + 8);
    a_.swap(b);
  }
  bool empty() const { return n_ == 0; }
  size_t size() const { return n_; }
  uint64_t top() const { return a_[0]; }

  // priority_queue::push: push_back, then __push_heap(first, n-1, 0, value)
  void push(uint64_t v) {
    if (a_.size() < 2 * (n_ + 1) + 8) a_.resize(2 * a_.size() + 2 * (n_ + 1) + 8);
    ++n_;
    sift_up((ptrdiff_t)n_ - 1, 0, v);
  }

  // priority_queue::pop: pop_heap (when n > 1: value = last, last = first,
  // __adjust_heap(first, 0, n-1, value)), then pop_back
  void pop() {
    const ptrdiff_t n = (ptrdiff_t)n_;
    if (n > 1) {
      uint64_t value = a_[n - 1];
      a_[n - 1] = a_[0];
      adjust(0, n - 1, value);
    }
    --n_;
  }

 private:
  // a_[0, n_) is the heap; the rest is zero-filled room for look-ahead loads
  std::vector<uint64_t> a_;
  size_t n_ = 0;

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

  // __adjust_heap: the hole walks down to a leaf, always to the child of
  // higher priority (the right one on ties), then the value sifts up.  Each
  // step of the loop covers two levels; the loads of the second level do not
  // depend on the first level's comparison.
  void adjust(ptrdiff_t hole, ptrdiff_t len, uint64_t value) {
    uint64_t *a = a_.data();
    const ptrdiff_t top = hole;
    const ptrdiff_t lim = (len - 1) / 2;  // nodes below lim have two children
    while (hole < lim) {
      const uint64_t l = a[2 * hole + 1], r = a[2 * hole + 2];
      const uint64_t *g = a + 4 * hole + 3;  // grandchildren, < 2 * len + 4
      const uint64_t g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
      const bool left = comp(r, l);
      a[hole] = left ? l : r;
      hole = 2 * hole + 2 - left;
      if (hole >= lim) break;
      const uint64_t cl = left ? g0 : g2, cr = left ? g1 : g3;
      const bool left2 = comp(cr, cl);
      a[hole] = left2 ? cl : cr;
      hole = 2 * hole + 2 - left2;
    }
    if ((len & 1) == 0 && hole == (len - 2) / 2) {
      const ptrdiff_t second = 2 * (hole + 1);
      a[hole] = a[second - 1];
      hole = second - 1;
    }
    sift_up(hole, top, value);
  }
};

}  // namespace mscx
