// Test-only: checks that mscx::ExactHeap pops elements in the same order as
// std::priority_queue with the reference's comparator semantics.
#include <cstdint>
#include <queue>
#include <vector>
#include "../mscomplex_amd/csrc/heap.hpp"

struct Elem { uint32_t key, payload; };
struct Cmp { bool operator()(const Elem &a, const Elem &b) const { return a.key != b.key ? a.key > b.key : false; } };

static uint64_t rng(uint64_t &s) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }

extern "C" long heap_compare(uint64_t seed, int nops, int nkeys, int push_pct) {
  std::priority_queue<Elem, std::vector<Elem>, Cmp> ref;
  mscx::ExactHeap mine;
  long mismatches = 0;
  uint32_t next = 0;
  for (int i = 0; i < nops; ++i) {
    bool push = ref.empty() || (int)(rng(seed) % 100) < push_pct;
    if (push) {
      uint32_t k = (uint32_t)(rng(seed) % (uint64_t)nkeys);
      ref.push(Elem{k, next});
      mine.push(mscx::ExactHeap::make(k, next));
      ++next;
    } else {
      Elem a = ref.top();
      uint64_t b = mine.top();
      if (a.payload != mscx::ExactHeap::payload(b) || a.key != mscx::ExactHeap::key(b)) ++mismatches;
      ref.pop();
      mine.pop();
    }
  }
  while (!ref.empty()) {
    if (ref.top().payload != mscx::ExactHeap::payload(mine.top())) ++mismatches;
    ref.pop();
    mine.pop();
  }
  return mismatches + (mine.empty() ? 0 : 1000000);
}
