// Per-wave body of the fs3d tick (shared by the single-world and batched kernels).
//
// One tick updates every 2x2x2 block of a grid (blocks shifted by (-p,-p,-p), p =
// tick parity).  A "row quad" is the 4 rows (y, y+1) x (z, z+1) of one block row:
// every block of a row quad reads and writes only those 4 rows, so row quads are
// independent, and because one wave loads all of a row quad before storing any of
// it, the update may also be done in place (src == dst).  Each lane owns a segment
// of S bytes of the 4 rows (lanes_per_row = ceil(nx / S) <= 64 lanes per row quad);
// a wave packs floor(64 / lanes_per_row) row quads.  Loads and stores are S-byte
// vector accesses, coalesced across the lanes of a row: 1 byte in and 1 byte out
// per voxel.
//
// Odd parity shifts the blocks by one byte against the lane segments: every lane
// takes the last byte of its left neighbour and the first byte of its right
// neighbour through full-wave DPP shifts (wave_shr:1 / wave_shl:1), evaluates
// S/2 + 1 blocks on the shifted frame, and keeps the outputs of its own bytes.
// The straddling block is evaluated by both lanes from identical inputs.
//
// Block evaluation is a table lookup.  The 4 rows' 2-bit categories of a block are
// packed into a 16-bit state with SWAR byte operations (two blocks per dword); the
// table (4 variants x 65536 states x 8 bytes, 2 MiB, cache-resident) returns the 8
// output categories as bytes, which v_perm scatters back into row dwords.
// Reserved bytes (>= 4) take a wave-uniform slower path that maps them to STATIC
// and restores their value afterwards.
#pragma once

#include "fs3d_internal.h"

namespace fs3d {
namespace tickcore {

typedef u32 u32x2 __attribute__((ext_vector_type(2)));
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
// Row starts are arbitrary byte offsets; gfx9 global accesses do not require
// natural alignment, so segments are moved with unaligned vector accesses.
typedef u32x2 __attribute__((aligned(1))) u32x2_u;
typedef u32x4 __attribute__((aligned(1))) u32x4_u;
typedef u64 u64x2 __attribute__((ext_vector_type(2)));
typedef u64x2 __attribute__((aligned(8))) u64x2_a8;  // two adjacent mask words

constexpr u32 kStatic4 = 0x01010101u;

__device__ __forceinline__ u32 lane_from_left(u32 v) {  // lane L receives lane L-1
  return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ u32 lane_from_right(u32 v) {  // lane L receives lane L+1
  return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, false);
}

// 0xFF in every byte whose value is >= 4 (reserved materials), else 0x00.
__device__ __forceinline__ u32 reserved_mask(u32 x) {
  const u32 nz = (((x & 0x7C7C7C7Cu) + 0x7C7C7C7Cu) | x) & 0x80808080u;
  return (nz - (nz >> 7)) | nz;
}

// Categories: byte values 0..3 unchanged, reserved bytes -> STATIC (1).
__device__ __forceinline__ u32 to_codes(u32 x, u32 rmask) {
  return (x & 0x03030303u & ~rmask) | (rmask & kStatic4);
}

__device__ __forceinline__ u32 fmix32(u32 h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ u64 dmix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Keeps the low `keep` bytes of x (0..4) and fills the others with STATIC.
__device__ __forceinline__ u32 keep_bytes(u32 x, int keep) {
  if (keep >= 4) return x;
  if (keep <= 0) return kStatic4;
  const u32 m = (1u << (8 * keep)) - 1u;
  return (x & m) | (kStatic4 & ~m);
}

// NT: streaming (nontemporal) accesses, for grids that do not stay in cache
// from one tick to the next.
template <int S, bool NT = false>
__device__ __forceinline__ void load_seg(const u8 *p, u32 (&d)[S / 4]) {
  if constexpr (S == 8) {
    const u32x2_u *q = reinterpret_cast<const u32x2_u *>(p);
    const u32x2 v = NT ? __builtin_nontemporal_load(q) : *q;
    d[0] = v.x;
    d[1] = v.y;
  } else {
#pragma unroll
    for (int c = 0; c < S / 16; c++) {
      const u32x4_u *q = reinterpret_cast<const u32x4_u *>(p + 16 * c);
      const u32x4 v = NT ? __builtin_nontemporal_load(q) : *q;
      d[4 * c + 0] = v.x;
      d[4 * c + 1] = v.y;
      d[4 * c + 2] = v.z;
      d[4 * c + 3] = v.w;
    }
  }
}

template <int S, bool NT = false>
__device__ __forceinline__ void store_seg(u8 *p, const u32 (&d)[S / 4]) {
  if constexpr (S == 8) {
    u32x2 v;
    v.x = d[0];
    v.y = d[1];
    if constexpr (NT)
      __builtin_nontemporal_store(v, reinterpret_cast<u32x2_u *>(p));
    else
      *reinterpret_cast<u32x2_u *>(p) = v;
  } else {
#pragma unroll
    for (int c = 0; c < S / 16; c++) {
      u32x4 v;
      v.x = d[4 * c + 0];
      v.y = d[4 * c + 1];
      v.z = d[4 * c + 2];
      v.w = d[4 * c + 3];
      if constexpr (NT)
        __builtin_nontemporal_store(v, reinterpret_cast<u32x4_u *>(p + 16 * c));
      else
        *reinterpret_cast<u32x4_u *>(p + 16 * c) = v;
    }
  }
}

// Packs the 16-bit states of the two blocks held by one frame dword of the 4
// rows: low half = block at bytes (0,1), high half = block at bytes (2,3).
__device__ __forceinline__ u32 pack_states(u32 c0, u32 c1, u32 c2, u32 c3) {
  const u32 r01 = c0 | (c1 << 4);
  const u32 r23 = c2 | (c3 << 4);
  const u32 a = r01 | (r01 >> 6);
  const u32 b = r23 | (r23 >> 6);
  return __builtin_amdgcn_perm(b, a, 0x06020400u);
}

// ---- settled-segment skipping ----------------------------------------------------
// A tick may skip a lane segment (no load, no store) when every block the lane
// evaluates is known not to change.  Whether a block state is a fixed point of the
// rule does not depend on the variant (a variant only picks among eligible
// targets, never whether one exists), so a block that did not change at tick t-2
// and none of whose cells changed at tick t-1 does not change at tick t; and the
// destination buffer, which received tick t-2's output, already holds its bytes.
// Changes are recorded per parity and row quad as a mask over lane segments: bit s
// is set when lane s evaluated a block that changed.  At tick t (parity p) a lane
// reads its own bit from the parity-p masks (written at tick t-2) and, of the four
// parity-(1-p) row quads that share its rows (tick t-1), bits s-1..s+1 on odd
// ticks and bit s on even ticks, then rewrites its row quad's parity-p mask.  Host-side edits set the bits of the cells they
// touch (or every bit), so edited cells are recomputed at both parities.
struct SkipCtx {
  u64 *mine;          // parity-p masks: read (tick t-2), then rewritten (tick t)
  const u64 *other;   // parity-(1-p) masks (tick t-1)
  int nyq_o, nzq_o;   // row-quad counts of parity 1-p
  bool load_skipped;  // skipped lanes still load (hash ticks); they never store
};

// One lane's share of one row quad: its S-byte segment of the 4 rows (r = dy +
// 2*dz), loaded, plus where it lives.  Out-of-grid rows read as STATIC.
// Off is the type of the rows' byte offsets: u32 for grids of at most 2^32
// bytes (one register per address, a 32-bit offset from the uniform base
// pointer), else size_t.  ST: which of the segment's loads (kStreamLoads) and
// stores (kStreamStores) are streaming (nontemporal) accesses.
constexpr int kStreamLoads = 1, kStreamStores = 2;
template <int S, class Off = size_t, int ST = 0>
struct RowQuad {
  u32 own[4][S / 4];
  Off rowoff[4];
  long long q;  // row-quad index within the tick
  int valid;    // bit r: row r is inside the grid
  int m;        // valid bytes of the segment
  int seg, yq, zq;
  bool active;  // the lane holds a row quad of this tick
  bool skip;    // settled: nothing to evaluate or store
  bool loaded;  // own[] holds the grid bytes (false: skipped without loading)
  // Packed tick only (rq_update's EDGES): the left / right neighbour segment is
  // settled and its straddling block keeps its input (nl / nr); or its edge
  // bytes are read from memory into edge[] (ml / mr; edge[r] = left byte << 24 |
  // right byte); sl / sr: the neighbour is not being updated, so this lane also
  // stores the neighbour's cell of the straddling block when it changes.
  bool nl, nr, ml, mr, sl, sr;
  u32 edge[4];
};

// Loads the lane's segment of the 4 rows of row quad (R.yq, R.zq) at parity p
// (R.seg and R.active set); fetch = false leaves placeholders (STATIC).
template <int S, class Off, int ST>
__device__ __forceinline__ void rq_fetch(RowQuad<S, Off, ST> &R, const u8 *src, int nx, int ny, int nz, int p,
                                         bool fetch) {
  constexpr int ND = S / 4;
  R.loaded = fetch;
  const int y0 = 2 * R.yq - p, z0 = 2 * R.zq - p;
  const int xs = R.seg * S;
  R.m = min(S, nx - xs);
  R.valid = 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int y = y0 + (r & 1), z = z0 + (r >> 1);
    const bool ok = R.active && y >= 0 && y < ny && z >= 0 && z < nz;
    R.valid |= ok ? (1 << r) : 0;
    R.rowoff[r] = (Off)(((size_t)(ok ? z : 0) * ny + (ok ? y : 0)) * nx + xs);
#pragma unroll
    for (int j = 0; j < ND; j++) R.own[r][j] = kStatic4;
    if (ok && fetch) load_seg<S, (ST & kStreamLoads) != 0>(src + R.rowoff[r], R.own[r]);
  }
  // a partial last segment: bytes past the row read as STATIC (after all four
  // loads are issued, so that they overlap)
  if (R.m < S) {
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int j = 0; j < ND; j++) R.own[r][j] = keep_bytes(R.own[r][j], R.m - 4 * j);
  }
}

// Sets R.skip for an active lane from the change masks (R.seg, R.yq, R.zq, R.q
// set).
// The three mask loads (the own word, and per other-parity z row the two
// adjacent words y = yq - p, yq + 1 - p as one 16-byte load) are issued
// unconditionally (out-of-range ones from word 0, then discarded) so that they
// overlap: loads under per-lane branches are each waited for before the next is
// issued.  A pair read past the last word of a parity's row quads reads its
// edge words, which follow in the same allocation (512^3 box scene: 45.3
// against 46.1 us per tick over ticks 5-25 with the last hashed, 35.8 against
// 36.7 over ticks 100-1100, with five 8-byte loads).
template <int S, class Off, int ST>
__device__ __forceinline__ void rq_skip_check(RowQuad<S, Off, ST> &R, int p, const SkipCtx *sk) {
  u64x2 v[2];
  bool in[4];  // word (y = yq + (k & 1) - p, z = zq + (k >> 1) - p) is a row quad
  long long i0[2];
#pragma unroll
  for (int kz = 0; kz < 2; kz++) {
    const int zo = R.zq + kz - p;
    const bool zin = R.active && zo >= 0 && zo < sk->nzq_o;
#pragma unroll
    for (int ky = 0; ky < 2; ky++) {
      const int yo = R.yq + ky - p;
      in[2 * kz + ky] = zin && yo >= 0 && yo < sk->nyq_o;
    }
    // a pair that would start before word 0 (y = -1 in row 0) is read from
    // word 0, which is then its second word
    i0[kz] = zin ? (long long)zo * sk->nyq_o + R.yq - p : 0;
    v[kz] = *reinterpret_cast<const u64x2_a8 *>(sk->other + (i0[kz] < 0 ? 0 : i0[kz]));
  }
  const u64 mine = sk->mine[R.active ? R.q : 0];
  u64 o = 0;
#pragma unroll
  for (int kz = 0; kz < 2; kz++) {
    const u64 w0 = v[kz].x, w1 = i0[kz] < 0 ? v[kz].x : v[kz].y;
    o |= (in[2 * kz] ? w0 : 0ull) | (in[2 * kz + 1] ? w1 : 0ull);
  }
  // Odd ticks: segments seg, seg+1 and seg-1 at tick t-1 (the straddling
  // blocks read a byte of each neighbour; shifted right only: seg may be 63).
  // Even ticks: segment seg alone.  No block straddles a segment edge at even
  // parity, and every block that touched one of this segment's cells at the
  // odd tick t-1 was evaluated by this segment's lane there, so its bit holds
  // every change to them.
  const u64 near = p ? ((o >> R.seg) & 3u) | (R.seg > 0 ? (o >> (R.seg - 1)) & 1u : 0u) : (o >> R.seg) & 1u;
  const u64 own = (mine >> R.seg) & 1u;  // this segment at tick t-2
  R.skip = R.active && (near | own) == 0;
}

// Loads row-quad group `wave` (in units of rqpw row quads) of a tick of parity p.
template <int S, bool SKIP = false>
__device__ __forceinline__ void rq_load(RowQuad<S> &R, const u8 *src, int nx, int ny, int nz, int LR, int rqpw,
                                        int p, long long wave, int lane, const SkipCtx *sk = nullptr) {
  const int nyq = (ny + p + 1) >> 1, nzq = (nz + p + 1) >> 1;
  const long long nrq = (long long)nyq * nzq;
  const int rql = lane / LR;
  R.seg = lane - rql * LR;
  const long long q = wave * rqpw + rql;
  const bool active = rql < rqpw && q < nrq;
  R.yq = active ? (int)(q % nyq) : 0;
  R.zq = active ? (int)(q / nyq) : 0;
  R.q = q;
  R.active = active;
  R.skip = false;
  if constexpr (SKIP) rq_skip_check(R, p, sk);
  rq_fetch<S>(R, src, nx, ny, nz, p, !SKIP || !R.skip || sk->load_skipped);
}

// Evaluates the blocks of a loaded row quad at tick t and stores the new bytes.
// Returns this lane's share of the hash of the written cells when HASH.  Every
// lane of the wave must call it together (DPP exchange).
// With SKIP or TRACK, *changed reports whether a block this lane evaluated
// changed.
// With EDGES (segment-list tick), a lane's neighbours come from R.nl / nr / ml /
// mr / edge instead of the adjacent lanes' load state.
// P = t & 1 when known at compile time (even ticks need no neighbour exchange,
// no shifted frame and no extra frame dword), or -1: the parity is read from t.
// With EDGES, *edge_changed receives bit 0 / bit 1 when this lane stored the
// left / right neighbour's cell (R.sl / R.sr) because it changed, and bit 2 /
// bit 3 when the lane's first / last block changed (on odd ticks the blocks
// straddling its left / right edge).
template <int S, bool HASH, bool SKIP, bool TRACK, bool EDGES, int P, class Off, int ST>
__device__ __forceinline__ u64 rq_update_p(const RowQuad<S, Off, ST> &R, u8 *dst, int LR, u32 t,
                                           const u64 *__restrict__ rule, bool *changed, u32 *edge_changed) {
  constexpr int ND = S / 4;  // dwords per lane per row
  const int p = P < 0 ? (int)(t & 1u) : P;
  const int seg = R.seg;

  // ---- reserved bytes: wave-uniform slow path ----
  // (edge bytes read from memory enter the frame too)
  u32 anyres = 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int j = 0; j < ND; j++) anyres |= R.own[r][j] & 0xFCFCFCFCu;
    if constexpr (EDGES) anyres |= R.edge[r] & 0xFC0000FCu;
  }
  const bool slow = __any(anyres != 0);

  // ---- frame: even parity = own bytes; odd parity = [left byte, own, right byte] ----
  constexpr int NF = ND + 1;  // frame dwords (the last is used by odd parity only)
  u32 fr[4][NF];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    u32 from_l = lane_from_left(R.own[r][ND - 1]);
    u32 from_r = lane_from_right(R.own[r][0]);
    if constexpr (EDGES) {  // the adjacent lane may hold an unrelated segment
      from_l = R.ml ? R.edge[r] : (R.nl ? kStatic4 : from_l);
      from_r = R.mr ? R.edge[r] : (R.nr ? kStatic4 : from_r);
    }
    if (p) {
      const u32 lb = seg == 0 ? 0x01000000u : (from_l & 0xFF000000u);
      // frame bytes 2-3 of the last dword hold no block: STATIC, like the outputs
      const u32 rb = (seg == LR - 1 ? 0x01u : (from_r & 0xFFu)) | 0x00010100u;
      fr[r][0] = __builtin_amdgcn_alignbyte(R.own[r][0], lb, 3);
#pragma unroll
      for (int j = 1; j < ND; j++) fr[r][j] = __builtin_amdgcn_alignbyte(R.own[r][j], R.own[r][j - 1], 3);
      fr[r][ND] = __builtin_amdgcn_alignbyte(rb, R.own[r][ND - 1], 3);
    } else {
#pragma unroll
      for (int j = 0; j < ND; j++) fr[r][j] = R.own[r][j];
      fr[r][ND] = kStatic4;
    }
  }

  // ---- neighbours skipped without loading (odd parity) ----
  // Their edge bytes are placeholders; the block straddling into such a
  // neighbour is settled (its own bit covered it), so it keeps its input.
  bool nl = false, nr = false;
  if constexpr (SKIP && EDGES) {
    nl = p && R.nl;
    nr = p && R.nr;
  } else if constexpr (SKIP) {
    const u32 unl = R.loaded ? 0u : 1u;
    const u32 fl = lane_from_left(unl), fr_ = lane_from_right(unl);
    nl = p && seg > 0 && fl != 0;
    nr = p && seg < LR - 1 && fr_ != 0;
  }
  u32 dif = 0;
  u32 ends = 0;  // EDGES: bit 0 / 1: the first / last block changed

  // ---- variants: one 32-bit word per 16 consecutive blocks ----
  // A lane evaluates blocks bx0 .. bx0 + S/2 (the last one on odd ticks only);
  // with bx0 & 15 = jb they span groups g0 .. g0 + (jb + S/2) / 16.
  constexpr int NG = S >= 64 ? 3 : 2;
  const u32 bx0 = (u32)(seg * (S / 2));
  const u32 g0 = bx0 >> 4;
  const u32 yzt = (u32)R.yq * 0x85EBCA77u + (u32)R.zq * 0xC2B2AE3Du + t * 0x27D4EB2Fu;
  u32 wg[NG];
#pragma unroll
  for (int i = 0; i < NG; i++) wg[i] = fmix32((g0 + (u32)i) * 0x9E3779B1u + yzt);
  const u32 jb = (S >= 32) ? 0u : (bx0 & 15u);
  // S < 32: the lane's S/2 + 1 <= 9 blocks start at block jb of group g0, so one
  // 32-bit word (the two groups' words shifted by 2 * jb) holds all their
  // variants at compile-time offsets.
  const u32 vlane = (S < 32) ? (u32)(((((u64)wg[1]) << 32) | wg[0]) >> (2u * jb)) : 0u;

  // Reserved bytes read as STATIC: the frame becomes categories in one
  // (rarely taken) uniform branch, so that the evaluation below is one basic
  // block in which the scheduler can issue every table gather before the first
  // wait.
  if (slow) {
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int j = 0; j < NF; j++) fr[r][j] = to_codes(fr[r][j], reserved_mask(fr[r][j]));
  }

  // ---- evaluate blocks (two per frame dword) and scatter outputs ----
  u32 out[4][NF];
#pragma unroll
  for (int j = 0; j < NF; j++) {
    if (j == ND && !p) {
#pragma unroll
      for (int r = 0; r < 4; r++) out[r][j] = kStatic4;
      continue;
    }
    const u32 states = pack_states(fr[0][j], fr[1][j], fr[2][j], fr[3][j]);
    u64 e[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (j == ND && h == 1) {  // odd parity's last frame dword holds one block
        e[1] = 0x0101010101010101ull;
        continue;
      }
      u32 v;
      if constexpr (S < 32) {
        v = (vlane >> (2 * (2 * j + h))) & 3u;
      } else {  // jb = 0: the group and the offset are compile-time
        constexpr int kb = 0;
        const int k = kb + 2 * j + h;
        v = (wg[k >> 4] >> (2 * (k & 15))) & 3u;
      }
      const u32 s = h ? (states >> 16) : (states & 0xFFFFu);
      e[h] = rule[(v << 16) | s];
    }
    const u32 lo0 = (u32)e[0], hi0 = (u32)(e[0] >> 32), lo1 = (u32)e[1], hi1 = (u32)(e[1] >> 32);
    // Entries flag a changing block in bit 7 of cell 0's byte (out[0] bytes 0, 2).
    bool keep_in = false;
    if constexpr (SKIP) keep_in = (j == 0 && nl) || (j == ND && nr);  // block 0 straddles into a settled lane
    if constexpr (SKIP || TRACK) dif |= (keep_in ? lo1 : (lo0 | lo1)) & (u32)kRuleChanged;
    if constexpr (EDGES) {
      // the first block is block 0 of dword 0; the last is block 0 of the extra
      // dword on odd ticks and block 1 of dword ND - 1 on even ones
      if (j == 0) ends |= ((keep_in ? 0u : lo0) & (u32)kRuleChanged) ? 1u : 0u;
      if (j == (p ? ND : ND - 1)) ends |= ((p ? (keep_in ? 0u : lo0) : lo1) & (u32)kRuleChanged) ? 2u : 0u;
    }
    out[0][j] = __builtin_amdgcn_perm(lo1, lo0, 0x05040100u) & 0x7F7F7F7Fu;
    out[1][j] = __builtin_amdgcn_perm(lo1, lo0, 0x07060302u);
    out[2][j] = __builtin_amdgcn_perm(hi1, hi0, 0x05040100u);
    out[3][j] = __builtin_amdgcn_perm(hi1, hi0, 0x07060302u);
    if constexpr (SKIP) {
      // The settled straddling block keeps its input: of its bytes, only this
      // lane's own (frame byte 1 at j = 0, frame byte 0 at j = ND) is stored.
      // Selects, not a branch, keep the loop one basic block.
      if (j == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++)
          out[r][0] = keep_in ? (out[r][0] & 0xFFFF00FFu) | ((R.own[r][0] & 0xFFu) << 8) : out[r][0];
      } else if (j == ND) {
#pragma unroll
        for (int r = 0; r < 4; r++)
          out[r][ND] = keep_in ? (out[r][ND] & 0xFFFFFF00u) | (R.own[r][ND - 1] >> 24) : out[r][ND];
      }
    }
  }
  if constexpr (SKIP || TRACK) {
    if (changed) *changed = R.loaded && dif != 0;
  }
  // The neighbour's cell of a straddling block whose other lane is not updated
  // this tick: stored here when it changed (categories 0..3; a reserved byte
  // never moves).
  if constexpr (EDGES) {
    u32 echg = 0;
    if (p) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (R.valid >> r & 1) {
          const u32 old_l = R.edge[r] >> 24, new_l = out[r][0] & 0xFFu;
          if (R.sl && old_l < 4u && new_l != old_l) {
            dst[R.rowoff[r] - 1] = (u8)new_l;
            echg |= 1u;
          }
          const u32 old_r = R.edge[r] & 0xFFu, new_r = (out[r][ND] >> 8) & 0xFFu;
          if (R.sr && old_r < 4u && new_r != old_r) {
            dst[R.rowoff[r] + S] = (u8)new_r;
            echg |= 2u;
          }
        }
      }
    }
    if (edge_changed) *edge_changed = echg | (ends << 2);
  }

  // ---- back to own bytes, restore reserved values, store, hash ----
  u64 acc = 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    u32 res[ND];
#pragma unroll
    for (int j = 0; j < ND; j++) {
      u32 o = p ? __builtin_amdgcn_alignbyte(out[r][j + 1], out[r][j], 1) : out[r][j];
      if (slow) {
        const u32 rm = reserved_mask(R.own[r][j]);
        o = (R.own[r][j] & rm) | (o & ~rm);
      }
      res[j] = o;
    }
    if (R.valid >> r & 1) {
      u8 *dp = dst + R.rowoff[r];
      if (SKIP && R.skip) {
        // settled: the destination already holds these bytes
      } else if (R.m == S) {
        store_seg<S, (ST & kStreamStores) != 0>(dp, res);
      } else {
#pragma unroll
        for (int k = 0; k < S; k++)
          if (k < R.m) dp[k] = (u8)(res[k >> 2] >> (8 * (k & 3)));
      }
      if constexpr (HASH) {
        // Each non-empty cell costs a 64-bit mix; a dword position that is
        // empty in every lane of the wave is skipped as a whole (a uniform
        // branch), which spares most of the work in empty space.
#pragma unroll
        for (int j = 0; j < ND; j++) {
          if (__any(res[j] != 0u)) {
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
              const int k = 4 * j + kk;
              const u32 v = (res[j] >> (8 * kk)) & 0xFFu;
              if (k < R.m && v) acc += dmix64((((u64)R.rowoff[r] + (u64)k) << 8) | v);
            }
          }
        }
      }
    }
  }
  return acc;
}

// One copy of the update per parity, chosen by a branch that is uniform: the
// tick's parity is the same across a wave in every caller (kernel arguments, or
// one world per workgroup).
template <int S, bool HASH, bool SKIP = false, bool TRACK = SKIP, bool EDGES = false, class Off = size_t,
          int ST = 0>
__device__ __forceinline__ u64 rq_update(const RowQuad<S, Off, ST> &R, u8 *dst, int LR, u32 t,
                                         const u64 *__restrict__ rule, bool *changed = nullptr,
                                         u32 *edge_changed = nullptr) {
  if (t & 1u) return rq_update_p<S, HASH, SKIP, TRACK, EDGES, 1>(R, dst, LR, t, rule, changed, edge_changed);
  return rq_update_p<S, HASH, SKIP, TRACK, EDGES, 0>(R, dst, LR, t, rule, changed, edge_changed);
}

// A single copy for both parities, for the batched kernels' persistent loops
// (two copies there double the live registers).
template <int S, bool HASH, bool SKIP = false, bool TRACK = SKIP>
__device__ __forceinline__ u64 rq_update_any(const RowQuad<S> &R, u8 *dst, int LR, u32 t,
                                             const u64 *__restrict__ rule, bool *changed = nullptr) {
  return rq_update_p<S, HASH, SKIP, TRACK, false, -1>(R, dst, LR, t, rule, changed, nullptr);
}

// Processes row-quad group `wave` of one tick at tick counter t (src may equal
// dst).  Every lane of the wave must call it.
// With TRACK, *changed reports whether a block of this lane changed.
template <int S, bool HASH, bool TRACK = false>
__device__ __forceinline__ u64 tick_wave(const u8 *src, u8 *dst, int nx, int ny, int nz, int LR, int rqpw, u32 t,
                                         long long wave, int lane, const u64 *__restrict__ rule,
                                         bool *changed = nullptr) {
  RowQuad<S> R;
  rq_load<S>(R, src, nx, ny, nz, LR, rqpw, (int)(t & 1u), wave, lane);
  return rq_update_any<S, HASH, false, TRACK>(R, dst, LR, t, rule, changed);
}

// Processes NR row-quad groups (wave0, wave0 + stride, ...) with every load
// issued before the first update, so NR groups' memory latency overlaps.  Row
// quads are disjoint, so loading ahead is safe also in place.  Groups past the
// end of the tick are inactive (no loads or stores).
// With `changed`, ORs into *changed whether any block of this lane changed.
template <int S, bool HASH, int NR>
__device__ __forceinline__ u64 tick_waves(const u8 *src, u8 *dst, int nx, int ny, int nz, int LR, int rqpw, u32 t,
                                          long long wave0, long long stride, int lane, const u64 *__restrict__ rule,
                                          bool *changed = nullptr) {
  RowQuad<S> R[NR];
  const int p = (int)(t & 1u);
#pragma unroll
  for (int i = 0; i < NR; i++) rq_load<S>(R[i], src, nx, ny, nz, LR, rqpw, p, wave0 + i * stride, lane);
  u64 acc = 0;
#pragma unroll
  for (int i = 0; i < NR; i++) {
    bool ch = false;
    acc += rq_update_any<S, HASH, false, true>(R[i], dst, LR, t, rule, &ch);
    if (changed) *changed = *changed || ch;
  }
  return acc;
}

}  // namespace tickcore
}  // namespace fs3d
