// Block rule table of the fs3d tick (host side).
//
// The rule is this project's own (the reference has no simulation: its only
// per-frame state update is camera motion, src/engine/rendering/renderer.cpp:437-467).
// A 2x2x2 block holds cells b = ox + 2*oy + 4*oz (oz = 1 on top), each with a
// category EMPTY 0, STATIC 1, SAND 2, FLUID 3.  The four columns (ox + 2*oy) are
// walked on the ring [0, 1, 3, 2]; variant v scans ring positions starting at v in
// direction +1 (v even) or -1 (v odd), and a cell at ring position k prefers
// targets at k+dir, k-dir, then k+2.  Every cell swaps at most once per tick:
//   fall   - per column: SAND/FLUID over EMPTY swap; SAND over FLUID swap;
//   slide  - unmoved top SAND -> first unmoved EMPTY/FLUID bottom cell (other columns);
//   topple - unmoved top FLUID -> first unmoved EMPTY bottom cell (other columns);
//   spread - bottom layer then top layer: unmoved FLUID -> first unmoved EMPTY
//            cell of another column in the same layer.
// This file evaluates the rule with per-category bit masks (bit b = cell b) and
// tabulates it for all 4^8 block states and 4 variants.
#include "fs3d_internal.h"

namespace fs3d {

namespace {

constexpr int kRing[4] = {0, 1, 3, 2};

struct Masks {
  u32 cat[4];   // bit b set when cell b has that category
  u32 moved;
  int get(int b) const {
    for (int c = 0; c < 4; c++)
      if (cat[c] >> b & 1u) return c;
    return 1;
  }
  void swap(int a, int b) {
    int ca = get(a), cb = get(b);
    cat[ca] ^= (1u << a);
    cat[cb] ^= (1u << b);
    cat[ca] |= (1u << b);
    cat[cb] |= (1u << a);
    moved |= (1u << a) | (1u << b);
  }
  bool free_of(int b, u32 want) const {  // unmoved and category in `want` (bit per category)
    if (moved >> b & 1u) return false;
    for (int c = 0; c < 4; c++)
      if ((want >> c & 1u) && (cat[c] >> b & 1u)) return true;
    return false;
  }
};

// Moves the unmoved cell at ring position k of `layer_src` (0 bottom, 1 top) to the
// first acceptable target in preference order on `layer_dst`; returns true on a move.
bool move_first(Masks &m, int k, int dir, int layer_src, int layer_dst, u32 want) {
  const int off[3] = {dir, -dir, 2};
  const int src = kRing[k] + 4 * layer_src;
  for (int i = 0; i < 3; i++) {
    const int kk = (k + off[i] + 8) & 3;
    const int dst = kRing[kk] + 4 * layer_dst;
    if (m.free_of(dst, want)) {
      m.swap(src, dst);
      return true;
    }
  }
  return false;
}

Masks evaluate(u16 s, int v) {
  Masks m{};
  for (int b = 0; b < 8; b++) m.cat[(s >> (2 * b)) & 3] |= 1u << b;
  const int dir = (v & 1) ? -1 : 1;
  const u32 bit_empty = 1u << FS3D_EMPTY, bit_fluid = 1u << FS3D_FLUID;

  // fall: columns whose top is movable over an empty bottom, or sand over fluid
  const u32 top_sand = (m.cat[FS3D_SAND] >> 4) & 0xFu, top_fluid = (m.cat[FS3D_FLUID] >> 4) & 0xFu;
  const u32 bot_empty = m.cat[FS3D_EMPTY] & 0xFu, bot_fluid = m.cat[FS3D_FLUID] & 0xFu;
  const u32 falls = ((top_sand | top_fluid) & bot_empty) | (top_sand & bot_fluid);
  for (int col = 0; col < 4; col++)
    if (falls >> col & 1u) m.swap(col + 4, col);

  // slide, topple: top cells in scan order
  for (int pass = 0; pass < 2; pass++) {
    const int want_src = pass == 0 ? FS3D_SAND : FS3D_FLUID;
    const u32 want_dst = pass == 0 ? (bit_empty | bit_fluid) : bit_empty;
    for (int i = 0; i < 4; i++) {
      const int k = (v + dir * i + 8) & 3;
      const int top = kRing[k] + 4;
      if ((m.moved >> top & 1u) || !(m.cat[want_src] >> top & 1u)) continue;
      move_first(m, k, dir, 1, 0, want_dst);
    }
  }
  // spread within a layer, bottom first
  for (int layer = 0; layer < 2; layer++)
    for (int i = 0; i < 4; i++) {
      const int k = (v + dir * i + 8) & 3;
      const int c = kRing[k] + 4 * layer;
      if ((m.moved >> c & 1u) || !(m.cat[FS3D_FLUID] >> c & 1u)) continue;
      move_first(m, k, dir, layer, layer, bit_empty);
    }
  return m;
}

}  // namespace

u16 rule_state_packed(u16 s, int v) {
  Masks m = evaluate(s, v);
  u16 r = 0;
  for (int b = 0; b < 8; b++) r |= (u16)(m.get(b) << (2 * b));
  return r;
}

void build_rule_entries(u64 *entries) {
  for (int v = 0; v < 4; v++)
    for (u32 s = 0; s < 65536u; s++) {
      Masks m = evaluate((u16)s, v);
      u64 e = 0;
      u32 r = 0;
      for (int b = 0; b < 8; b++) {
        e |= (u64)m.get(b) << (8 * b);
        r |= (u32)m.get(b) << (2 * b);
      }
      if (r != s) e |= kRuleChanged;
      entries[(size_t)v * 65536u + s] = e;
    }
}

}  // namespace fs3d
