// GridArena: the serving loop's device memory for resident grids, carved out of
// a few large chunks, with the budget enforced on the bytes actually allocated
// (chunks), not on the bytes handed out.
//
// Blocks are 256-byte aligned, placed best fit, split when the rest is worth
// keeping, and coalesced with free neighbours of the same chunk when released,
// so any free space serves any size.  A chunk is allocated only while the
// allocated total stays within the budget; when a block does not fit, chunks
// with nothing live in them are released first.  A request that still does not
// fit gets kOverBudget (the caller waits for live blocks to come back), and a
// failed chunk allocation gets kNoMemory.  Host-only bookkeeping: the chunk
// allocator is a callback, so the arena is tested on the CPU with host memory.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace fs3d {

class GridArena {
 public:
  enum Result { kOk = 0, kOverBudget = 1, kNoMemory = 2 };
  typedef void *(*ChunkAlloc)(size_t bytes, void *ctx);  // nullptr on failure
  typedef void (*ChunkFree)(void *p, void *ctx);
  static constexpr size_t kAlign = 256;
  static constexpr size_t kMinSplit = 4096;  // smaller remainders stay with the block

  GridArena() = default;
  GridArena(const GridArena &) = delete;
  GridArena &operator=(const GridArena &) = delete;
  ~GridArena() { clear(); }

  void init(size_t budget, size_t chunk, ChunkAlloc a, ChunkFree f, void *ctx) {
    budget_ = budget;
    chunk_ = chunk;
    alloc_ = a;
    free_ = f;
    ctx_ = ctx;
  }

  // Frees every chunk (live blocks included).
  void clear() {
    for (auto &c : chunks_)
      if (free_) free_(c.first, ctx_);
    chunks_.clear();
    free_by_addr_.clear();
    free_by_size_.clear();
    live_.clear();
    allocated_ = in_use_ = 0;
  }

  Result alloc(size_t bytes, uint8_t **out) {
    const size_t b = round_up(bytes ? bytes : 1);
    if (b > budget_) return kOverBudget;
    for (int attempt = 0; attempt < 2; attempt++) {
      if (take(b, out)) return kOk;
      // no free block fits: a new chunk, within the budget
      size_t cs = b > chunk_ ? b : chunk_;
      if (allocated_ + cs > budget_) cs = b;
      if (allocated_ + cs > budget_) {
        if (attempt == 0 && release_idle_chunks()) continue;
        return kOverBudget;
      }
      uint8_t *p = static_cast<uint8_t *>(alloc_(cs, ctx_));
      if (!p) return kNoMemory;
      chunks_[p] = Chunk{cs, 0};
      allocated_ += cs;
      if (allocated_ > peak_allocated_) peak_allocated_ = allocated_;
      insert_free(p, cs);
      if (take(b, out)) return kOk;
      return kNoMemory;  // unreachable: the new chunk holds b bytes
    }
    return kOverBudget;
  }

  // Returns a block from alloc; unknown pointers are ignored.
  void release(uint8_t *p) {
    auto it = live_.find(p);
    if (it == live_.end()) return;
    size_t size = it->second;
    live_.erase(it);
    auto ch = chunk_of(p);
    ch->second.used -= size;
    in_use_ -= size;
    uint8_t *const lo = ch->first, *const hi = ch->first + ch->second.size;
    // coalesce with the following and the preceding free block of this chunk
    auto nx = free_by_addr_.find(p + size);
    if (nx != free_by_addr_.end() && nx->first < hi) {
      size += nx->second;
      erase_free(nx);
    }
    auto pv = free_by_addr_.lower_bound(p);
    if (pv != free_by_addr_.begin()) {
      --pv;
      if (pv->first >= lo && pv->first + pv->second == p) {
        p = pv->first;
        size += pv->second;
        erase_free(pv);
      }
    }
    insert_free(p, size);
  }

  size_t allocated() const { return allocated_; }
  size_t peak_allocated() const { return peak_allocated_; }
  size_t in_use() const { return in_use_; }
  size_t peak_in_use() const { return peak_in_use_; }
  size_t budget() const { return budget_; }
  size_t chunks() const { return chunks_.size(); }
  size_t live_blocks() const { return live_.size(); }
  size_t free_blocks() const { return free_by_addr_.size(); }

  static size_t round_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

 private:
  struct Chunk {
    size_t size;
    size_t used;
  };

  std::map<uint8_t *, Chunk>::iterator chunk_of(uint8_t *p) { return std::prev(chunks_.upper_bound(p)); }

  void insert_free(uint8_t *p, size_t size) {
    free_by_addr_[p] = size;
    free_by_size_.emplace(size, p);
  }

  void erase_free(std::map<uint8_t *, size_t>::iterator it) {
    auto r = free_by_size_.equal_range(it->second);
    for (auto s = r.first; s != r.second; ++s)
      if (s->second == it->first) {
        free_by_size_.erase(s);
        break;
      }
    free_by_addr_.erase(it);
  }

  // Best fit among the free blocks; splits off a remainder of kMinSplit or more.
  bool take(size_t b, uint8_t **out) {
    auto s = free_by_size_.lower_bound(b);
    if (s == free_by_size_.end()) return false;
    uint8_t *p = s->second;
    size_t size = s->first;
    erase_free(free_by_addr_.find(p));
    if (size - b >= kMinSplit) {
      insert_free(p + b, size - b);
      size = b;
    }
    live_[p] = size;
    chunk_of(p)->second.used += size;
    in_use_ += size;
    if (in_use_ > peak_in_use_) peak_in_use_ = in_use_;
    *out = p;
    return true;
  }

  // Frees the chunks with no live block; true when any was freed.
  bool release_idle_chunks() {
    bool any = false;
    for (auto it = chunks_.begin(); it != chunks_.end();) {
      auto f = free_by_addr_.find(it->first);
      // an idle chunk is one free block (releases coalesce within a chunk)
      if (it->second.used == 0 && f != free_by_addr_.end() && f->second == it->second.size) {
        erase_free(f);
        allocated_ -= it->second.size;
        free_(it->first, ctx_);
        it = chunks_.erase(it);
        any = true;
      } else {
        ++it;
      }
    }
    return any;
  }

  size_t budget_ = 0, chunk_ = 0;
  ChunkAlloc alloc_ = nullptr;
  ChunkFree free_ = nullptr;
  void *ctx_ = nullptr;
  std::map<uint8_t *, Chunk> chunks_;
  std::map<uint8_t *, size_t> free_by_addr_;
  std::multimap<size_t, uint8_t *> free_by_size_;
  std::map<uint8_t *, size_t> live_;
  size_t allocated_ = 0, peak_allocated_ = 0, in_use_ = 0, peak_in_use_ = 0;
};

}  // namespace fs3d
