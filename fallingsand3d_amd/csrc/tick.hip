// fs3d tick kernels for gfx950 (CDNA4, wave64).
//
// One tick updates every 2x2x2 block of the grid (blocks shifted by (-p,-p,-p),
// p = tick parity) from the source buffer into the destination buffer.  The
// reference has no tick; its engine loop (src/engine/engine.cpp:59-69) is the
// call site this replaces.
//
// The per-wave work (row quads, DPP exchange, table lookup) lives in
// tick_core.h; this file holds the single-world launches, the hash passes and
// small helper kernels.
#include "tick_core.h"

namespace fs3d {

namespace {

using namespace tickcore;

struct TickArgs {
  const u8 *src;
  u8 *dst;
  int nx, ny, nz;
  int lanes_per_row, rq_per_wave;
  u32 lr_magic;  // ceil(65536 / lanes_per_row): lane / lanes_per_row = (lane * lr_magic) >> 16
  u64 t;
  u32 k_off;
  const u64 *tick_dev;  // when set, tick = *tick_dev + k_off (graph replays)
  const u64 *rule;      // 4 x 65536 entries
  u64 *parts;           // hash partials (HASH only)
  u64 *masks[2];        // per-parity change masks (SKIP only)
  size_t edge_off;      // edge words of the masks: masks[p] + edge_off (SKIP only)
  u64 *stats;           // optional sweep counters (SKIP only): kHashParts slots + violations
  u32 *ctl;             // work-list counts, one per tick mod 3 (SKIP only)
  int dilate;           // plan pass: odd ticks also list the neighbours of changed segments
};

__device__ __forceinline__ u64 tick_of(const TickArgs &a) { return a.tick_dev ? *a.tick_dev + a.k_off : a.t; }

// Sector size of the full-grid skipping tick's load and store decisions: a
// partly written 64-byte sector costs more than a whole one (512^3 box scene,
// ticks 5-25 with the last hashed: 50.1 against 52.0 us per tick with
// segment-sized decisions; ticks 100-300: 41.2 against 42.2).
constexpr int kSkipSectorBytes = 64;

// The grid is x: groups of 4 * rq_per_wave row quads along y (one group per
// workgroup), y and z: z row quads (zq = blockIdx.z * 65535 + blockIdx.y), so a
// lane's row quad comes from the block indices without a division.  Off: the
// row offsets' type (u32 when the grid has at most 2^32 bytes).
template <int S, bool HASH, bool SKIP, class Off>
__global__ __launch_bounds__(256) void tick_kernel(TickArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const u64 t64 = tick_of(a);
  const u32 t = (u32)t64;
  const int p = (int)(t & 1u);
  const int nyq = (a.ny + p + 1) >> 1, nzq = (a.nz + p + 1) >> 1;
  const int LR = a.lanes_per_row, rqpw = a.rq_per_wave;
  const int rql = (int)(((u32)lane * a.lr_magic) >> 16);
  const int seg = lane - rql * LR;
  const int yq = ((int)blockIdx.x * 4 + wv) * rqpw + rql;
  const int zq = (int)blockIdx.z * 65535 + (int)blockIdx.y;
  const long long bid = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const long long wave = bid * 4 + wv;
  // A hash tick replaces the plan pass of the listed tick: it clears the count
  // slot the next tick's plan pass appends to.
  if (SKIP && HASH && a.ctl && bid == 0 && threadIdx.x == 0) a.ctl[(t64 + 1) % 3] = 0;
  SkipCtx sk;
  if constexpr (SKIP) {
    sk.mine = a.masks[p];
    sk.other = a.masks[p ^ 1];
    sk.nyq_o = (a.ny + (p ^ 1) + 1) >> 1;
    sk.nzq_o = (a.nz + (p ^ 1) + 1) >> 1;
    sk.load_skipped = HASH;  // hash ticks read settled segments to hash them
  }
  const u64 lane_bits = LR == 64 ? ~0ull : ((1ull << LR) - 1ull);
  // Whole waves past the end of the row leave early; the hashing variant keeps
  // them until its workgroup barrier (they load and store nothing).
  const bool active = rql < rqpw && yq < nyq && zq < nzq;
  if (!HASH && (__builtin_amdgcn_readfirstlane((int)(((int)blockIdx.x * 4 + wv) * rqpw)) >= nyq || zq >= nzq))
    return;
  RowQuad<S, Off> R;
  R.seg = seg;
  R.active = active;
  R.yq = active ? yq : 0;
  R.zq = active ? zq : 0;
  R.q = (long long)R.zq * nyq + R.yq;
  R.skip = false;
  bool rule_skip = false;  // settled by the skip rule (before sector merging)
  if constexpr (SKIP) {
    rq_skip_check(R, p, &sk);
    rule_skip = R.skip;
    // A settled segment whose 64-byte sector holds an active one is evaluated
    // too (it keeps its bytes), so that stores cover whole sectors.  A sector
    // is an aligned group of G lanes; of it, the lanes of this lane's row quad
    // count (every lane of a row quad shares its active flag).
    if constexpr (S < kSkipSectorBytes) {
      constexpr int G = kSkipSectorBytes / S;
      const u64 act = __ballot(R.active && !R.skip);
      const int g0 = lane & ~(G - 1), r0 = rql * LR;
      const int lo = max(g0, r0), hi = min(g0 + G, r0 + LR);
      const u64 group = hi > lo ? ((1ull << (hi - lo)) - 1ull) << lo : 0ull;
      R.skip = R.skip && (act & group) == 0;
    }
  }
  rq_fetch<S>(R, a.src, a.nx, a.ny, a.nz, p, !SKIP || !R.skip || sk.load_skipped);
  if constexpr (SKIP && !HASH) {
    if (__ballot(R.active && !R.skip) == 0) {  // the whole wave is settled
      if (R.active && R.seg == 0) {
        sk.mine[R.q] = 0;
        sk.mine[a.edge_off + R.q] = 0;
      }
      return;
    }
  }
  bool ch = false;
  u64 acc = rq_update<S, HASH, SKIP>(R, a.dst, LR, t, a.rule, &ch);
  if constexpr (SKIP) {
    const u64 b = __ballot(ch);
    if (R.active && R.seg == 0) {
      // edge bits: a changed segment's blocks touch its own edges only
      const u64 m = (b >> (lane - R.seg)) & lane_bits;
      sk.mine[R.q] = m;
      sk.mine[a.edge_off + R.q] = (m | (m << 1)) & lane_bits;
    }
    if (a.stats) {
      const u64 vox = R.active ? (u64)R.m * (u64)__builtin_popcount(R.valid) : 0;
      // swept: the segments the skip rule evaluates (sector merging adds
      // traffic, not work the rule needs); read-only: settled segments read
      // for the hash
      const u64 swept = wave_sum_u64(rule_skip ? 0 : vox);
      const u64 ronly = wave_sum_u64(rule_skip && HASH ? vox : 0);
      const u64 bad = __ballot(rule_skip && ch);  // a settled segment changed: must never happen
      if (lane == 0) {
        u64 *slot = a.stats + (size_t)(wave % kHashParts) * kHashStride;
        atomicAdd((unsigned long long *)slot, (unsigned long long)swept);
        if (ronly) atomicAdd((unsigned long long *)(slot + 1), (unsigned long long)ronly);
        if (bad)
          atomicAdd((unsigned long long *)(a.stats + (size_t)kHashParts * kHashStride),
                    (unsigned long long)__builtin_popcountll(bad));
      }
    }
  }
  if constexpr (HASH) {
    __shared__ u64 wsum[4];
    acc = wave_sum_u64(acc);
    if (lane == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      const u64 tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
      if (tot) atomicAdd((unsigned long long *)(a.parts + (size_t)(bid % kHashParts) * kHashStride),
                         (unsigned long long)tot);
    }
  }
}

// ---- listed (work-list) skipping tick ------------------------------------------
// A plan pass turns the change masks into the lane-activity word of every row
// quad, (mine | o | o << 1 | o >> 1) over the 64 lane bits on odd ticks and
// (mine | o) on even ticks, o being the OR of the four other-parity masks that
// share its rows; row quads with no active lane
// are settled (their mask is already 0 and stays so) and cost nothing more, the
// others are appended to a work list.  A persistent kernel then streams the list.

constexpr int kPlanPerThread = 4;  // row quads per plan thread: 1024 per workgroup

// Appends the active row quads of a 1024-row-quad range to the list, in row-quad
// order within the range, with one atomic per workgroup on the count slot of
// this tick (tick mod 3); clears the slot of the next tick.
__global__ __launch_bounds__(256) void plan_kernel(TickArgs a, TickEntry *list, u32 *ctl) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const u64 t64 = tick_of(a);
  const int p = (int)(t64 & 1u);
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl[(t64 + 1) % 3] = 0;
  const int nyq = (a.ny + p + 1) >> 1, nzq = (a.nz + p + 1) >> 1;
  const int nyq_o = (a.ny + (p ^ 1) + 1) >> 1, nzq_o = (a.nz + (p ^ 1) + 1) >> 1;
  const long long nrq = (long long)nyq * nzq;
  const u64 lane_bits = a.lanes_per_row == 64 ? ~0ull : ((1ull << a.lanes_per_row) - 1ull);
  const u64 *other = a.masks[p ^ 1];
  // odd ticks of the packed schedule: which segment edges may have a changing
  // straddling block (edge words of this parity's row quad, two ticks before,
  // and of the four other-parity row quads sharing its rows, one tick before)
  const bool edges = p && !a.dilate;
  u64 act[kPlanPerThread], bal[kPlanPerThread], edg[kPlanPerThread];
  int yq[kPlanPerThread], zq[kPlanPerThread];
  __shared__ u32 wcount[kPlanPerThread * 4];
  __shared__ u32 wbase;
#pragma unroll
  for (int i = 0; i < kPlanPerThread; i++) {
    const long long q = (long long)blockIdx.x * (256 * kPlanPerThread) + i * 256 + threadIdx.x;
    act[i] = 0;
    edg[i] = 0;
    yq[i] = zq[i] = 0;
    if (q < nrq) {
      yq[i] = (int)(q % nyq);
      zq[i] = (int)(q / nyq);
      // unconditional loads (out-of-range ones from word 0, discarded) overlap
      u64 v[4];
      bool in[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int yo = yq[i] + (k & 1) - p, zo = zq[i] + (k >> 1) - p;
        in[k] = yo >= 0 && yo < nyq_o && zo >= 0 && zo < nzq_o;
        v[k] = other[in[k] ? (long long)zo * nyq_o + yo : 0];
      }
      const u64 own = a.masks[p][q];
      u64 o = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) o |= in[k] ? v[k] : 0ull;
      if (edges) {
        u64 ev[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int yo = yq[i] + (k & 1) - p, zo = zq[i] + (k >> 1) - p;
          ev[k] = other[a.edge_off + (in[k] ? (long long)zo * nyq_o + yo : 0)];
        }
        u64 ed = a.masks[p][a.edge_off + q];
#pragma unroll
        for (int k = 0; k < 4; k++) ed |= in[k] ? ev[k] : 0ull;
        edg[i] = ed;
      }
      // The neighbours' bits matter on odd ticks only (straddling blocks), and
      // only for the listed schedule: the packed one evaluates a straddling
      // block next to an unlisted segment from its true bytes instead.
      act[i] = (own | o | (p && a.dilate ? (o << 1) | (o >> 1) : 0ull)) & lane_bits;
    }
    bal[i] = __ballot(act[i] != 0);
    if (lane == 0) wcount[i * 4 + wv] = (u32)__builtin_popcountll(bal[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 tot = 0;
    for (int k = 0; k < kPlanPerThread * 4; k++) {  // exclusive prefix in (i, wave) order
      const u32 c = wcount[k];
      wcount[k] = tot;
      tot += c;
    }
    wbase = tot ? atomicAdd(&ctl[t64 % 3], tot) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPlanPerThread; i++) {
    if (act[i]) {
      const u32 idx = wbase + wcount[i * 4 + wv] + (u32)__builtin_popcountll(bal[i] & ((1ull << lane) - 1ull));
      list[idx].act = act[i];
      list[idx].pos = ((u64)(u32)zq[i] << 32) | (u32)yq[i];
      if (!a.dilate) list[idx].edge = edg[i];
    }
  }
}

template <int S>
__global__ __launch_bounds__(256) void tick_list_kernel(TickArgs a, const TickEntry *__restrict__ list, u32 *ctl,
                                                        u32 nwaves) {
  const int lane = threadIdx.x & 63;
  const u32 wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const u64 t64 = tick_of(a);
  const u32 t = (u32)t64;
  const int p = (int)(t & 1u);
  const u32 count = ctl[t64 % 3];
  const int LR = a.lanes_per_row, rqpw = a.rq_per_wave;
  const int rql = lane / LR, seg = lane - rql * LR;
  const int nyq = (a.ny + p + 1) >> 1;
  const u64 lane_bits = LR == 64 ? ~0ull : ((1ull << LR) - 1ull);
  u64 *mine = a.masks[p];
  const u64 stride = (u64)nwaves * rqpw;
  // Entries are prefetched one iteration ahead, so an item's chain is its data
  // loads, the update and the stores, as in the one-row-quad-per-wave kernel.
  u64 act_n = 0, pos_n = 0;
  {
    const u64 it = (u64)wave * rqpw + rql;
    if (rql < rqpw && it < count) {
      act_n = list[it].act;
      pos_n = list[it].pos;
    }
  }
  for (u64 base = (u64)wave * rqpw; base < count; base += stride) {
    const bool active = rql < rqpw && base + rql < count;
    const u64 act = act_n, pos = pos_n;
    act_n = 0;
    pos_n = 0;
    const u64 itn = base + stride + rql;
    if (rql < rqpw && itn < count) {
      act_n = list[itn].act;
      pos_n = list[itn].pos;
    }
    RowQuad<S> R;
    // An opaque copy of the (loop-invariant) segment index: otherwise every
    // per-lane predicate derived from it (byte-count masks of partial segments,
    // edge tests) is hoisted out of the loop and held in SGPR pairs, which spill.
    int seg_it = seg;
    asm volatile("" : "+v"(seg_it));
    R.seg = seg_it;
    R.active = active;
    R.yq = (int)(u32)pos;
    R.zq = (int)(pos >> 32);
    R.q = (long long)R.zq * nyq + R.yq;
    R.skip = active && !((act >> seg) & 1u);
    rq_fetch<S>(R, a.src, a.nx, a.ny, a.nz, p, !R.skip);
    bool ch = false;
    rq_update<S, false, true>(R, a.dst, LR, t, a.rule, &ch);
    const u64 b = __ballot(ch);
    if (active && seg == 0) {
      const u64 m = (b >> (lane - seg)) & lane_bits;
      mine[R.q] = m;
      mine[a.edge_off + R.q] = (m | (m << 1)) & lane_bits;
    }
    if (a.stats) {
      const u64 vox = active && !R.skip ? (u64)R.m * (u64)__builtin_popcount(R.valid) : 0;
      const u64 swept = wave_sum_u64(vox);
      if (lane == 0 && swept)
        atomicAdd((unsigned long long *)(a.stats + (size_t)(wave % kHashParts) * kHashStride),
                  (unsigned long long)swept);
    }
  }
}

// ---- packed skipping tick -------------------------------------------------------
// The listed tick gives each listed row quad a whole wave, which idles most lanes
// when activity is sparse within rows, and a wave pays for the block update
// whatever its active lanes.  This kernel reads the same list (same plan pass),
// a chunk of entries per wave, and packs their active 64-byte sectors into full
// waves: a sector of the four rows is one item of G = 64 / S lanes (S < 64), each
// lane one S-byte segment of it, so every item loads and stores whole 64-byte
// sectors.  HBM moves lines of 128 bytes for reads and sectors of 64 bytes for
// writes: a store of part of a sector costs more than a store of all of it (at
// 8 % of the 32-byte segments of a 2048^3 grid, scattered over its row quads,
// writing their whole sectors took 1.14 ms against 1.50 ms for the segments
// alone), and a sector's other segment, settled or not, comes with the same
// line.  Item i of the chunk is the k-th active sector of entry o, with o found
// by a binary search over the lanes' inclusive prefix of popcounts and the
// sector by a select of the k-th set bit.  Items are in (entry, sector) order,
// so a lane's left neighbour segment is being updated exactly when it is in the
// same item or the entry's sector bit to the left is set, and then it sits in
// the previous lane (or, for lane 0, in the previous group).
// On odd ticks a straddling block always sees its true bytes: the neighbour's
// edge bytes come from the adjacent lane (DPP) when it holds that segment, and
// from memory otherwise.  The plan pass therefore lists no neighbours of changed
// segments (unlike the listed schedule), and a lane next to a sector that is not
// listed stores that sector's cell of the straddling block itself when it
// changes, setting the segment's change bit.  A settled segment evaluated with
// its sector keeps its bytes and its change bit stays clear.  Change bits are
// ORed per entry within a group (segmented scan) and routed to the entry's
// lane, which writes the entry's mask word after the last group.

// Bit j: one of the G segment bits G*j .. G*j + G - 1 of act is set.
template <int G>
__device__ __forceinline__ u64 sector_bits(u64 act) {
  if constexpr (G == 1) {
    return act;
  } else {
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < 64 / G; j++) s |= (u64)(((act >> (G * j)) & ((1ull << G) - 1ull)) != 0ull) << j;
    return s;
  }
}

template <int S, int G>
__global__ __launch_bounds__(256) void tick_pack_kernel(TickArgs a, const TickEntry *__restrict__ list,
                                                        const u32 *ctl, u32 nwaves, int chunk) {
  static_assert(G >= 1 && 64 % G == 0, "lanes per item divide the wave");
  constexpr int kItems = 64 / G;  // items per group of 64 lanes
  const int lane = threadIdx.x & 63;
  const int h = lane & (G - 1);  // segment of the lane within its item
  const u32 wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const u64 t64 = tick_of(a);
  const u32 t = (u32)t64;
  const int p = (int)(t & 1u);
  const u64 count = ctl[t64 % 3];
  const int LR = a.lanes_per_row;
  const int nyq = (a.ny + p + 1) >> 1;
  const u64 lane_bits = LR == 64 ? ~0ull : ((1ull << LR) - 1ull);
  u64 *mine = a.masks[p];
  for (u64 c0 = (u64)wave * chunk; c0 < count; c0 += (u64)nwaves * chunk) {
    // this lane's entry of the chunk, and the prefix of active sectors
    const u64 e = c0 + lane;
    const bool has = lane < chunk && e < count;
    const u64 act = has ? list[e].act : 0ull;
    const u64 pos = has ? list[e].pos : 0ull;
    const u64 edge = (has && p) ? list[e].edge : 0ull;
    const u64 sec = sector_bits<G>(act);
    const int n = __builtin_popcountll(sec);
    int incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    const int excl = incl - n;
    const int total = __shfl(incl, 63, 64);
    u64 acc = 0, acc_e = 0;  // change and edge bits of this lane's entry
    for (int g0 = 0; g0 < total; g0 += kItems) {
      // An opaque copy of the item index: otherwise predicates derived from the
      // lane index are hoisted out of the loop into SGPR pairs.
      int idx = g0 + lane / G;
      asm volatile("" : "+v"(idx));
      const bool active = idx < total;
      // owner: the number of entries whose inclusive prefix is <= idx
      int o = 0;
#pragma unroll
      for (int w = 32; w >= 1; w >>= 1) {
        const int v = __shfl(incl, o + w - 1, 64);
        if (v <= idx) o += w;
      }
      o = active ? o : 63;
      int k = idx - __shfl(excl, o, 64);
      const u64 ao = ((u64)(u32)__shfl((int)(sec >> 32), o, 64) << 32) | (u32)__shfl((int)(u32)sec, o, 64);
      const u32 yq = (u32)__shfl((int)(u32)pos, o, 64), zq = (u32)__shfl((int)(pos >> 32), o, 64);
      u64 eo = 0;
      if (p) eo = ((u64)(u32)__shfl((int)(edge >> 32), o, 64) << 32) | (u32)__shfl((int)(u32)edge, o, 64);
      // the k-th set bit of ao: the item's sector
      int sj = 0;
#pragma unroll
      for (int w = 32; w >= 1; w >>= 1) {
        const int c = __builtin_popcountll((ao >> sj) & ((1ull << w) - 1ull));
        if (k >= c) {
          k -= c;
          sj += w;
        }
      }
      const int seg0 = sj * G + h;
      const bool lv = active && seg0 < LR;  // the lane holds a segment of the row
      const int seg = lv ? seg0 : 0;
      // neighbour segments being updated in this group or another one
      const bool left_act = lv && seg > 0 && (h > 0 || ((ao >> (sj - 1)) & 1ull));
      const bool right_act = lv && seg < LR - 1 && (h < G - 1 || ((ao >> (sj + 1)) & 1ull));
      // streaming stores: the active sectors of a large grid leave cache before the
      // next tick reads them (2048^3 pillar scene: 0.786 against 0.806 ms per tick;
      // streaming loads as well: 0.827)
      RowQuad<S, size_t, kStreamStores> R;
      R.seg = seg;
      R.active = lv;
      R.yq = (int)yq;
      R.zq = (int)zq;
      R.q = (long long)zq * nyq + yq;
      R.skip = false;
      // On odd ticks a straddling block sees its true bytes: from the adjacent
      // lane when it holds the neighbour segment, from memory otherwise (and
      // next to a segment that is not being updated this lane also stores the
      // neighbour's cell of that block).  A straddling block known not to
      // change (its edge bit is clear) keeps its input instead: nothing is read
      // from the neighbour, whose line then often stays unfetched.
      const bool lfix = p && !((eo >> seg) & 1ull);
      const bool rfix = p && seg < 63 && !((eo >> (seg + 1)) & 1ull);
      const bool lmem = lv && seg > 0 && (lane == 0 || !left_act);
      const bool rmem = lv && seg < LR - 1 && (lane == 63 || !right_act);
      R.nl = lmem && lfix;
      R.nr = rmem && rfix;
      R.ml = lmem && !lfix;
      R.mr = rmem && !rfix;
      R.sl = lv && seg > 0 && !left_act && !lfix;
      R.sr = lv && seg < LR - 1 && !right_act && !rfix;
      rq_fetch<S>(R, a.src, a.nx, a.ny, a.nz, p, true);
#pragma unroll
      for (int r = 0; r < 4; r++) R.edge[r] = 0;
      if (p && (R.ml || R.mr)) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          if (R.valid >> r & 1) {
            const u8 *row = a.src + R.rowoff[r];
            const u32 l = R.ml ? row[-1] : 0u;
            const u32 rt = R.mr ? row[S] : 0u;
            R.edge[r] = (l << 24) | rt;
          } else {
            R.edge[r] = kStatic4;
          }
        }
      }
      bool ch = false;
      u32 echg = 0;
      rq_update<S, false, true, true, true>(R, a.dst, LR, t, a.rule, &ch, &echg);
      // OR of the change bits of each entry's run of lanes in this group ...
      // A straddling block that changed sets the bit of a neighbour that is not
      // being updated too, even when only this lane's half of it moved: the
      // full-grid skip rule (hash ticks) reads a segment's bit as covering every
      // block it evaluates, the two straddling its edges included.
      u64 bits = (lv && ch) ? (1ull << seg) : 0ull;
      if (p && R.sl && (echg & 5u)) bits |= 1ull << (seg - 1);
      if (p && R.sr && (echg & 10u)) bits |= 1ull << (seg + 1);
      // edge bits: the lane's first block touches edge seg, its last edge seg + 1
      u64 ebits = 0;
      if (lv && (echg & 4u)) ebits |= 1ull << seg;
      if (lv && (echg & 8u) && seg < 63) ebits |= 1ull << (seg + 1);
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const u64 ob = (u64)__shfl_up((unsigned long long)bits, d, 64);
        const u64 oe = (u64)__shfl_up((unsigned long long)ebits, d, 64);
        const int oo = __shfl_up(o, d, 64);
        if (lane >= d && oo == o) {
          bits |= ob;
          ebits |= oe;
        }
      }
      // ... read by the entry's lane from the run's last lane
      const int s0 = max(excl, g0), s1 = min(incl, g0 + kItems);
      const int src = s1 > s0 ? (s1 - g0) * G - 1 : lane;
      const u64 run = (u64)__shfl((unsigned long long)bits, src, 64);
      const u64 run_e = (u64)__shfl((unsigned long long)ebits, src, 64);
      if (s1 > s0) {
        acc |= run;
        acc_e |= run_e;
      }
      if (a.stats) {
        // the listed segments only (a sector's settled segments add traffic,
        // not work the skip rule needs)
        const u64 aw = ((u64)(u32)__shfl((int)(act >> 32), o, 64) << 32) | (u32)__shfl((int)(u32)act, o, 64);
        const bool listed = lv && ((aw >> seg) & 1ull);
        const u64 vox = listed ? (u64)R.m * (u64)__builtin_popcount(R.valid) : 0;
        const u64 swept = wave_sum_u64(vox);
        if (lane == 0 && swept)
          atomicAdd((unsigned long long *)(a.stats + (size_t)(wave % kHashParts) * kHashStride),
                    (unsigned long long)swept);
      }
    }
    if (has) {
      const long long q = (long long)(u32)(pos >> 32) * nyq + (u32)pos;
      mine[q] = acc & lane_bits;
      mine[a.edge_off + q] = acc_e & lane_bits;
    }
  }
}

// ---- chunked dense tick ----------------------------------------------------------
// The dense tick with one wave per x-chunk of 64 * S bytes of a row quad's four
// rows, instead of one wave per whole rows: each wave moves fewer bytes (2 KiB
// in and out at S = 8) and more waves share the rows.  A chunk's end lanes
// read the neighbouring chunk's edge bytes from memory on odd ticks and
// evaluate the straddling block from the same inputs as the neighbouring
// chunk's end lane, each storing its own bytes.  Wave w of the grid takes
// chunk w % chunks of row quad w / chunks along y.
template <int S, bool HASH>
__global__ __launch_bounds__(256) void tick_chunk_kernel(TickArgs a, int chunks) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const u64 t64 = tick_of(a);
  const u32 t = (u32)t64;
  const int p = (int)(t & 1u);
  const int nyq = (a.ny + p + 1) >> 1, nzq = (a.nz + p + 1) >> 1;
  const int LR = a.lanes_per_row;  // segments per row (any number)
  const int w = (int)blockIdx.x * 4 + wv;
  const int chunk = w % chunks, yq = w / chunks;
  const int zq = (int)blockIdx.z * 65535 + (int)blockIdx.y;
  const long long bid = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  if (!HASH && (yq >= nyq || zq >= nzq)) return;  // uniform per wave
  const int seg = chunk * 64 + lane;
  const bool active = seg < LR && yq < nyq && zq < nzq;
  // stores stream (the grid does not stay in cache from one tick to the next);
  // loads stay cached for the chunk ends' edge bytes (2048^3 pillar scene, ms per
  // tick: 3.05 streaming stores, 3.12 streaming loads and stores, 3.20 neither)
  RowQuad<S, size_t, kStreamStores> R;
  R.seg = active ? seg : 0;
  R.active = active;
  R.yq = active ? yq : 0;
  R.zq = active ? zq : 0;
  R.q = 0;
  R.skip = false;
  R.nl = R.nr = R.sl = R.sr = false;
  R.ml = active && p && lane == 0 && seg > 0;
  R.mr = active && p && lane == 63 && seg < LR - 1;
  rq_fetch<S>(R, a.src, a.nx, a.ny, a.nz, p, true);
#pragma unroll
  for (int r = 0; r < 4; r++) R.edge[r] = kStatic4;
  if (R.ml || R.mr) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (R.valid >> r & 1) {
        const u8 *row = a.src + R.rowoff[r];
        const u32 l = R.ml ? row[-1] : 0u;
        const u32 rt = R.mr ? row[S] : 0u;
        R.edge[r] = (l << 24) | rt;
      }
    }
  }
  u64 acc = rq_update<S, HASH, false, false, true>(R, a.dst, LR, t, a.rule);
  if constexpr (HASH) {
    __shared__ u64 wsum[4];
    acc = wave_sum_u64(acc);
    if (lane == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      const u64 tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
      if (tot) atomicAdd((unsigned long long *)(a.parts + (size_t)(bid % kHashParts) * kHashStride),
                         (unsigned long long)tot);
    }
  }
}

// Full-grid state hash.  A wave takes 1 KiB (16 bytes per lane), compacts the
// non-empty cells into LDS (one 32-bit item per cell: value << 16 | byte offset
// within the 1 KiB) and mixes them 64 at a time, so each cell costs a mix only
// when it is non-empty; mixing under per-byte branches would cost the wave a
// mix for every byte position that any of its lanes has non-empty.
__global__ __launch_bounds__(256) void hash_kernel(const u8 *g, size_t n, u64 *parts) {
  __shared__ u32 items[4][1024];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  u32 *it = items[wv];
  u64 acc = 0;
  const size_t nchunk = (n + 15) / 16;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  // block-uniform loop (two barriers per pass); c0 is this wave's first chunk
  for (size_t cb = (size_t)blockIdx.x * blockDim.x; cb < nchunk; cb += stride) {
    const size_t c0 = cb + (size_t)wv * 64, c = c0 + lane;
    const size_t base = c * 16;
    u32 d[4] = {0u, 0u, 0u, 0u};
    if (c < nchunk) {
      const u32x4 v = *reinterpret_cast<const u32x4 *>(g + base);  // buffers are 256-byte aligned
      d[0] = v.x;
      d[1] = v.y;
      d[2] = v.z;
      d[3] = v.w;
    }
    u32 nz = 0;
#pragma unroll
    for (int k = 0; k < 16; k++)
      nz |= (((d[k >> 2] >> (8 * (k & 3))) & 0xFFu) != 0u && base + k < n) ? (1u << k) : 0u;
    const int cnt = __builtin_popcount(nz);
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    const int total = __shfl(incl, 63, 64);
    const int excl = incl - cnt;
#pragma unroll
    for (int k = 0; k < 16; k++)
      if ((nz >> k) & 1u)
        it[excl + __builtin_popcount(nz & ((1u << k) - 1u))] =
            (((d[k >> 2] >> (8 * (k & 3))) & 0xFFu) << 16) | (u32)(lane * 16 + k);
    __syncthreads();
    for (int j = lane; j < total; j += 64) {
      const u32 x = it[j];
      acc += dmix64(((u64)(c0 * 16 + (x & 0x3FFu)) << 8) | (x >> 16));
    }
    __syncthreads();
  }
  __shared__ u64 wsum[4];
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u64 tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (tot) atomicAdd((unsigned long long *)(parts + (size_t)(blockIdx.x % kHashParts) * kHashStride),
                       (unsigned long long)tot);
  }
}

__global__ __launch_bounds__(256) void hash_finalize_kernel(u64 *parts, u64 *out) {
  u64 acc = 0;
  for (int i = threadIdx.x; i < kHashParts; i += 256) {
    acc += parts[(size_t)i * kHashStride];
    parts[(size_t)i * kHashStride] = 0;
  }
  __shared__ u64 wsum[4];
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void set_u64_kernel(u64 *p, u64 v) {
  if (threadIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(256) void counts_kernel(const u8 *g, size_t n, u64 *counts) {
  __shared__ u32 h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    atomicAdd(&h[g[i]], 1u);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd((unsigned long long *)&counts[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

template <int S, class Off>
hipError_t launch_so(const dim3 &grid, const TickArgs &a, bool hash, bool skip, hipStream_t s) {
  if (hash && skip)
    hipLaunchKernelGGL((tick_kernel<S, true, true, Off>), grid, dim3(256), 0, s, a);
  else if (hash)
    hipLaunchKernelGGL((tick_kernel<S, true, false, Off>), grid, dim3(256), 0, s, a);
  else if (skip)
    hipLaunchKernelGGL((tick_kernel<S, false, true, Off>), grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((tick_kernel<S, false, false, Off>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

// 32-bit row offsets free registers: the 16-byte skipping kernel holds 60
// instead of 67 VGPRs, eight waves per SIMD instead of seven (512^3 box scene,
// with the ballot sector merge: ticks 5-25 with the last hashed 46.3 against
// 47.3 us per tick, ticks 100-1100 36.7 against 37.5).
template <int S>
hipError_t launch_s(const TickShape &sh, const TickArgs &a, bool hash, bool skip, hipStream_t s) {
  // sized for the larger parity's row-quad counts (graph replays read the tick
  // from the device); the kernel bounds itself by the tick's own
  const int nyq_max = (a.ny + 2) >> 1, nzq_max = (a.nz + 2) >> 1;
  const int per_block = 4 * sh.rq_per_wave;
  const dim3 grid((unsigned)((nyq_max + per_block - 1) / per_block), (unsigned)std::min(nzq_max, 65535),
                  (unsigned)((nzq_max + 65534) / 65535));
  if ((u64)a.nx * (u64)a.ny * (u64)a.nz <= (1ull << 32)) return launch_so<S, u32>(grid, a, hash, skip, s);
  return launch_so<S, size_t>(grid, a, hash, skip, s);
}

__global__ __launch_bounds__(256) void fill_u64_kernel(u64 *p, size_t n, u64 v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void mark_cells_kernel(u64 *m0, u64 *m1, long long i0, long long i1, u64 bits, size_t edge_off) {
  if (threadIdx.x == 0) {
    m0[i0] |= bits;
    m1[i1] |= bits;
    m0[edge_off + i0] |= bits | (bits << 1);
    m1[edge_off + i1] |= bits | (bits << 1);
  }
}

}  // namespace

int plan_tick(int nx, int ny, int nz, int seg_req, TickShape *out) {
  int seg = seg_req;
  if (seg == 0) seg = nx <= 1024 ? 16 : (nx <= 2048 ? 32 : 64);
  if (seg != 8 && seg != 16 && seg != 32 && seg != 64)
    return set_error(FS3D_ERR_INVALID, "segment must be 8, 16, 32 or 64 bytes (got %d)", seg);
  const int lpr = (nx + seg - 1) / seg;
  if (lpr > 64)
    return set_error(FS3D_ERR_INVALID, "nx=%d needs %d lanes per row with %d-byte segments (max 64)", nx,
                     lpr, seg);
  TickShape sh;
  sh.seg = seg;
  sh.lanes_per_row = lpr;
  sh.rq_per_wave = 64 / lpr;
  const long long nrq = (long long)((ny + 2) >> 1) * ((nz + 2) >> 1);
  sh.waves = (nrq + sh.rq_per_wave - 1) / sh.rq_per_wave;
  const long long blocks = (sh.waves + 3) / 4;
  if (blocks > 0x7FFFFFFFLL) return set_error(FS3D_ERR_INVALID, "grid too large");
  sh.blocks = (unsigned)blocks;
  *out = sh;
  return FS3D_OK;
}

hipError_t launch_tick(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t,
                       const u64 *tick_dev, u32 k_off, const u64 *rule, u64 *hash_parts,
                       hipStream_t s, u64 *const *masks, u64 *stats, u32 *ctl) {
  TickArgs a;
  a.src = src;
  a.dst = dst;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.lanes_per_row = sh.lanes_per_row;
  a.rq_per_wave = sh.rq_per_wave;
  a.lr_magic = (65536u + (u32)sh.lanes_per_row - 1u) / (u32)sh.lanes_per_row;
  a.t = t;
  a.k_off = k_off;
  a.tick_dev = tick_dev;
  a.rule = rule;
  a.parts = hash_parts;
  a.masks[0] = masks ? masks[0] : nullptr;
  a.masks[1] = masks ? masks[1] : nullptr;
  a.edge_off = mask_words_of(ny, nz);
  a.stats = masks ? stats : nullptr;
  a.ctl = masks ? ctl : nullptr;
  a.dilate = 1;
  const bool hash = hash_parts != nullptr;
  const bool skip = masks != nullptr;
  switch (sh.seg) {
    case 8: return launch_s<8>(sh, a, hash, skip, s);
    case 16: return launch_s<16>(sh, a, hash, skip, s);
    case 32: return launch_s<32>(sh, a, hash, skip, s);
    default: return launch_s<64>(sh, a, hash, skip, s);
  }
}

hipError_t launch_tick_chunked(int seg, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t, const u64 *tick_dev,
                               u32 k_off, const u64 *rule, u64 *hash_parts, hipStream_t s) {
  if (seg != 8 && seg != 16 && seg != 32) return hipErrorInvalidValue;
  TickArgs a;
  a.src = src;
  a.dst = dst;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.lanes_per_row = (nx + seg - 1) / seg;
  a.rq_per_wave = 1;
  a.lr_magic = 0;
  a.t = t;
  a.k_off = k_off;
  a.tick_dev = tick_dev;
  a.rule = rule;
  a.parts = hash_parts;
  a.masks[0] = a.masks[1] = nullptr;
  a.edge_off = 0;
  a.stats = nullptr;
  a.ctl = nullptr;
  a.dilate = 0;
  const int chunks = (a.lanes_per_row + 63) / 64;
  // sized for the larger parity's row-quad counts (the kernel bounds itself)
  const long long nyq_max = (ny + 2) >> 1, nzq_max = (nz + 2) >> 1;
  const long long waves = nyq_max * chunks;
  const dim3 grid((unsigned)((waves + 3) / 4), (unsigned)std::min<long long>(nzq_max, 65535),
                  (unsigned)((nzq_max + 65534) / 65535));
  const bool hash = hash_parts != nullptr;
  switch (seg) {
    case 8:
      if (hash) hipLaunchKernelGGL((tick_chunk_kernel<8, true>), grid, dim3(256), 0, s, a, chunks);
      else hipLaunchKernelGGL((tick_chunk_kernel<8, false>), grid, dim3(256), 0, s, a, chunks);
      break;
    case 16:
      if (hash) hipLaunchKernelGGL((tick_chunk_kernel<16, true>), grid, dim3(256), 0, s, a, chunks);
      else hipLaunchKernelGGL((tick_chunk_kernel<16, false>), grid, dim3(256), 0, s, a, chunks);
      break;
    default:
      if (hash) hipLaunchKernelGGL((tick_chunk_kernel<32, true>), grid, dim3(256), 0, s, a, chunks);
      else hipLaunchKernelGGL((tick_chunk_kernel<32, false>), grid, dim3(256), 0, s, a, chunks);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_tick_listed(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t,
                              const u64 *tick_dev, u32 k_off, const u64 *rule, hipStream_t s,
                              u64 *const *masks, u64 *stats, TickEntry *list, u32 *ctl, size_t mask_words,
                              u32 list_waves) {
  TickArgs a;
  a.src = src;
  a.dst = dst;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.lanes_per_row = sh.lanes_per_row;
  a.rq_per_wave = sh.rq_per_wave;
  a.lr_magic = (65536u + (u32)sh.lanes_per_row - 1u) / (u32)sh.lanes_per_row;
  a.t = t;
  a.k_off = k_off;
  a.tick_dev = tick_dev;
  a.rule = rule;
  a.parts = nullptr;
  a.masks[0] = masks[0];
  a.masks[1] = masks[1];
  a.edge_off = mask_words;
  a.stats = stats;
  a.ctl = ctl;
  a.dilate = 1;
  const size_t plan_blocks = (mask_words + 256 * kPlanPerThread - 1) / (256 * kPlanPerThread);
  hipLaunchKernelGGL(plan_kernel, dim3((unsigned)plan_blocks), dim3(256), 0, s, a, list, ctl);
  const dim3 grid(list_waves / 4);
  switch (sh.seg) {
    case 8: hipLaunchKernelGGL((tick_list_kernel<8>), grid, dim3(256), 0, s, a, list, ctl, list_waves); break;
    case 16: hipLaunchKernelGGL((tick_list_kernel<16>), grid, dim3(256), 0, s, a, list, ctl, list_waves); break;
    case 32: hipLaunchKernelGGL((tick_list_kernel<32>), grid, dim3(256), 0, s, a, list, ctl, list_waves); break;
    default: hipLaunchKernelGGL((tick_list_kernel<64>), grid, dim3(256), 0, s, a, list, ctl, list_waves); break;
  }
  return hipGetLastError();
}

// Lanes per item of the packed tick: one item per 64-byte sector of the rows
// for 16- and 32-byte segments (2048^3 pillar scene, 32-byte segments: 0.976
// against 1.092 ms per tick over ticks 0-60 with one lane per item).  8-byte
// segments keep one lane per item: eight lanes per item took 0.081 against
// 0.072 ms per tick on the 512^3 box scene (ticks 5-25), whose two buffers sit
// mostly in the last-level cache.
constexpr int kPackSectorBytes = 64;
constexpr int pack_g(int seg) { return seg >= kPackSectorBytes || seg < 16 ? 1 : kPackSectorBytes / seg; }

hipError_t launch_tick_packed(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t,
                              const u64 *tick_dev, u32 k_off, const u64 *rule, hipStream_t s, u64 *const *masks,
                              u64 *stats, TickEntry *list, u32 *ctl, size_t mask_words, u32 pack_waves,
                              int chunk) {
  TickArgs a;
  a.src = src;
  a.dst = dst;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.lanes_per_row = sh.lanes_per_row;
  a.rq_per_wave = sh.rq_per_wave;
  a.lr_magic = (65536u + (u32)sh.lanes_per_row - 1u) / (u32)sh.lanes_per_row;
  a.t = t;
  a.k_off = k_off;
  a.tick_dev = tick_dev;
  a.rule = rule;
  a.parts = nullptr;
  a.masks[0] = masks[0];
  a.masks[1] = masks[1];
  a.edge_off = mask_words;
  a.stats = stats;
  a.ctl = ctl;
  a.dilate = 0;
  const size_t plan_blocks = (mask_words + 256 * kPlanPerThread - 1) / (256 * kPlanPerThread);
  hipLaunchKernelGGL(plan_kernel, dim3((unsigned)plan_blocks), dim3(256), 0, s, a, list, ctl);
  const dim3 grid(pack_waves / 4);
  switch (sh.seg) {
    case 8: hipLaunchKernelGGL((tick_pack_kernel<8, pack_g(8)>), grid, dim3(256), 0, s, a, list, ctl, pack_waves, chunk); break;
    case 16: hipLaunchKernelGGL((tick_pack_kernel<16, pack_g(16)>), grid, dim3(256), 0, s, a, list, ctl, pack_waves, chunk); break;
    case 32: hipLaunchKernelGGL((tick_pack_kernel<32, pack_g(32)>), grid, dim3(256), 0, s, a, list, ctl, pack_waves, chunk); break;
    default: hipLaunchKernelGGL((tick_pack_kernel<64, pack_g(64)>), grid, dim3(256), 0, s, a, list, ctl, pack_waves, chunk); break;
  }
  return hipGetLastError();
}

int pack_blocks_per_cu(int seg) {
  int n = 0;
  hipError_t e;
  switch (seg) {
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_pack_kernel<8, pack_g(8)>, 256, 0); break;
    case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_pack_kernel<16, pack_g(16)>, 256, 0); break;
    case 32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_pack_kernel<32, pack_g(32)>, 256, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_pack_kernel<64, pack_g(64)>, 256, 0); break;
  }
  return (e == hipSuccess && n > 0) ? n : 1;
}

int list_blocks_per_cu(int seg) {
  int n = 0;
  hipError_t e;
  switch (seg) {
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_list_kernel<8>, 256, 0); break;
    case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_list_kernel<16>, 256, 0); break;
    case 32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_list_kernel<32>, 256, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tick_list_kernel<64>, 256, 0); break;
  }
  return (e == hipSuccess && n > 0) ? n : 1;
}

hipError_t launch_fill_u64(u64 *p, size_t n, u64 v, hipStream_t s) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_mark_cells(u64 *m0, u64 *m1, long long i0, long long i1, u64 bits, size_t edge_off, hipStream_t s) {
  hipLaunchKernelGGL(mark_cells_kernel, dim3(1), dim3(64), 0, s, m0, m1, i0, i1, bits, edge_off);
  return hipGetLastError();
}

hipError_t launch_hash(const u8 *g, size_t n, u64 *parts, hipStream_t s) {
  const size_t nchunk = (n + 15) / 16;
  size_t blocks = (nchunk + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(hash_kernel, dim3((unsigned)blocks), dim3(256), 0, s, g, n, parts);
  return hipGetLastError();
}

hipError_t launch_hash_finalize(u64 *parts, u64 *out, hipStream_t s) {
  hipLaunchKernelGGL(hash_finalize_kernel, dim3(1), dim3(256), 0, s, parts, out);
  return hipGetLastError();
}

hipError_t launch_set_u64(u64 *p, u64 v, hipStream_t s) {
  hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_counts(const u8 *g, size_t n, u64 *counts, hipStream_t s) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(counts_kernel, dim3((unsigned)blocks), dim3(256), 0, s, g, n, counts);
  return hipGetLastError();
}

}  // namespace fs3d
