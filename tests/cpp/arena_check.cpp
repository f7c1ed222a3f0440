// CPU check of the serving loop's grid arena (fallingsand3d_amd/csrc/arena.h)
// with host memory standing in for device chunks: random mixed-size streams of
// allocations and releases must never hold more chunk bytes than the budget,
// never hand out overlapping or out-of-chunk blocks, coalesce all free space of
// a chunk back into one block, and place any block up to the budget once
// everything has been released.  Prints "arena ok" and exits 0 on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "arena.h"

using fs3d::GridArena;

namespace {

struct HostChunks {
  std::map<uint8_t *, size_t> live;
  size_t fail_above = 0;  // allocations larger than this fail (0: never)
};

void *host_alloc(size_t bytes, void *ctx) {
  HostChunks *h = static_cast<HostChunks *>(ctx);
  if (h->fail_above && bytes > h->fail_above) return nullptr;
  uint8_t *p = static_cast<uint8_t *>(std::aligned_alloc(256, bytes));
  if (p) h->live[p] = bytes;
  return p;
}

void host_free(void *p, void *ctx) {
  HostChunks *h = static_cast<HostChunks *>(ctx);
  h->live.erase(static_cast<uint8_t *>(p));
  std::free(p);
}

int failures = 0;
#define CHECK(cond, ...)                          \
  do {                                            \
    if (!(cond)) {                                \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                   \
      std::printf("\n");                          \
      failures++;                                 \
      return;                                     \
    }                                             \
  } while (0)

// Blocks must lie inside one host chunk each and must not overlap.
void check_blocks(const HostChunks &h, const std::map<uint8_t *, size_t> &blocks) {
  uint8_t *end = nullptr;
  for (const auto &b : blocks) {
    CHECK(b.first >= end, "block %p overlaps the previous one", (void *)b.first);
    auto c = h.live.upper_bound(b.first);
    CHECK(c != h.live.begin(), "block %p below every chunk", (void *)b.first);
    --c;
    CHECK(b.first + b.second <= c->first + c->second, "block %p runs past its chunk", (void *)b.first);
    CHECK(((uintptr_t)b.first & 255) == 0, "block %p not 256-byte aligned", (void *)b.first);
    end = b.first + b.second;
  }
}

void random_stream(unsigned seed, size_t budget, size_t chunk) {
  HostChunks h;
  GridArena a;
  a.init(budget, chunk, host_alloc, host_free, &h);
  std::mt19937_64 rng(seed);
  std::map<uint8_t *, size_t> blocks;  // live block -> requested bytes
  std::vector<uint8_t *> order;
  size_t over = 0;
  for (int step = 0; step < 20000; step++) {
    const bool do_alloc = order.empty() || (rng() % 100) < 55;
    if (do_alloc) {
      // log-uniform sizes from 4 KiB to 1/3 of the budget
      const double lo = std::log(4096.0), hi = std::log((double)budget / 3);
      const size_t want = (size_t)std::exp(lo + (hi - lo) * (double)(rng() % 1000000) / 1e6);
      uint8_t *p = nullptr;
      const GridArena::Result r = a.alloc(want, &p);
      if (r == GridArena::kOk) {
        CHECK(blocks.find(p) == blocks.end(), "block %p handed out twice", (void *)p);
        blocks[p] = GridArena::round_up(want);
        order.push_back(p);
        std::fill(p, p + want, (uint8_t)step);  // the block is writable host memory
      } else {
        CHECK(r == GridArena::kOverBudget, "unexpected result %d", (int)r);
        over++;
      }
    } else {
      const size_t k = rng() % order.size();
      uint8_t *p = order[k];
      order[k] = order.back();
      order.pop_back();
      blocks.erase(p);
      a.release(p);
    }
    size_t held = 0;
    for (const auto &c : h.live) held += c.second;
    CHECK(held == a.allocated(), "arena reports %zu chunk bytes, %zu held", a.allocated(), held);
    CHECK(held <= budget, "chunk bytes %zu above the budget %zu", held, budget);
    if (step % 97 == 0) check_blocks(h, blocks);
    if (failures) return;
  }
  CHECK(a.peak_allocated() <= budget, "peak %zu above the budget", a.peak_allocated());
  CHECK(over > 0, "the stream never met the budget (budget %zu)", budget);
  for (uint8_t *p : order) a.release(p);
  CHECK(a.in_use() == 0 && a.live_blocks() == 0, "blocks left after releasing all");
  CHECK(a.free_blocks() == a.chunks(), "%zu free blocks in %zu chunks: free space not coalesced", a.free_blocks(),
        a.chunks());
  // everything free: the largest block the budget allows is placed (idle chunks go)
  uint8_t *p = nullptr;
  CHECK(a.alloc(budget, &p) == GridArena::kOk, "a budget-sized block does not fit an empty arena");
  CHECK(a.allocated() <= budget, "over budget after the large block");
  CHECK(a.alloc(1, &p) == GridArena::kOverBudget, "a block beyond a full budget was placed");
}

void device_refuses() {
  HostChunks h;
  h.fail_above = 1 << 20;
  GridArena a;
  a.init(64u << 20, 8u << 20, host_alloc, host_free, &h);
  uint8_t *p = nullptr;
  CHECK(a.alloc(4096, &p) == GridArena::kNoMemory, "a failed chunk allocation was not reported");
  CHECK(a.allocated() == 0 && a.chunks() == 0, "a failed chunk allocation left state behind");
  CHECK(a.alloc(100u << 20, &p) == GridArena::kOverBudget, "a block above the budget was not refused");
}

void small_blocks_share_a_chunk() {
  HostChunks h;
  GridArena a;
  a.init(64u << 20, 16u << 20, host_alloc, host_free, &h);
  uint8_t *p = nullptr, *q = nullptr;
  CHECK(a.alloc(4096, &p) == GridArena::kOk && a.alloc(5000, &q) == GridArena::kOk, "small blocks refused");
  CHECK(a.chunks() == 1 && q == p + 4096, "the second block did not follow the first in one chunk");
  CHECK(a.in_use() == 4096 + GridArena::round_up(5000), "in_use %zu", a.in_use());
}

}  // namespace

int main() {
  random_stream(1, 64u << 20, 16u << 20);
  random_stream(2, 40u << 20, 64u << 20);  // chunk larger than the budget
  random_stream(3, 200u << 20, 8u << 20);  // many small chunks
  device_refuses();
  small_blocks_share_a_chunk();
  if (failures) return 1;
  std::printf("arena ok\n");
  return 0;
}
