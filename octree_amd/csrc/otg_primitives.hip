// otg_primitives.hip -- scans and the LDS-ranked LSD radix sort used by the build.
//
// gfx950 specifics: 64-wide wavefronts; ranks within a wavefront come from
// ballot-based digit matching (8 ballots per round), per-wavefront digit
// counters live in LDS, and a 256-thread block (4 wavefronts) owns a tile of
// 2048 keys.
#include "otg_internal.hpp"

namespace otg {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;

// ---------------------------------------------------------------------------
// Generic block scan
// ---------------------------------------------------------------------------
// Count scans accumulate in 64 bits, so a total beyond 2^32 - 1 is seen (and
// reported saturated) instead of wrapping.
struct OpAdd64 {
  using T = uint64_t;
  __device__ static T identity() { return 0ull; }
  __device__ static T combine(T a, T b) { return a + b; }
  __device__ static T shfl_up(T x, int off) { return __shfl_up((unsigned long long)x, off, 64); }
};

// State of the "running minima" monoid over boundary levels.  For a segment,
// m = minimum level, mask = set of values taken by the running minimum when the
// segment is walked from its right end to its left end.  combine(left, right):
// walk right first, then continue into left where only values below right.m
// are new minima.
struct LevelState {
  uint32_t m;
  uint32_t mask;
};
struct OpLevel {
  using T = LevelState;
  __device__ static T identity() { return LevelState{255u, 0u}; }
  __device__ static uint32_t low_bits(uint32_t m) { return m >= 32u ? 0xFFFFFFFFu : ((1u << m) - 1u); }
  __device__ static T combine(T a, T b) {
    T r;
    r.m = a.m < b.m ? a.m : b.m;
    r.mask = b.mask | (a.mask & low_bits(b.m));
    return r;
  }
  __device__ static T shfl_up(T x, int off) {
    T r;
    r.m = __shfl_up(x.m, off, 64);
    r.mask = __shfl_up(x.mask, off, 64);
    return r;
  }
};

// Inclusive scan of one value per thread across the block.  Returns the
// inclusive value; *excl receives the exclusive value; *total the block total.
template <class Op>
__device__ typename Op::T block_scan(typename Op::T x, typename Op::T *excl, typename Op::T *total,
                                     typename Op::T *lds /* kBlock/64 */) {
  using T = typename Op::T;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  T inc = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T y = Op::shfl_up(inc, off);
    if (lane >= off) inc = Op::combine(y, inc);
  }
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  T wpre = Op::identity();
  T tot = Op::identity();
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) wpre = Op::combine(wpre, lds[w]);
    tot = Op::combine(tot, lds[w]);
  }
  __syncthreads();
  T prev = Op::shfl_up(inc, 1);
  T ex = lane == 0 ? wpre : Op::combine(wpre, prev);
  *excl = ex;
  *total = tot;
  return Op::combine(wpre, inc);
}

// ---- loaders / storers ----
struct U32Source {
  const uint32_t *p;
  __device__ uint64_t operator()(uint64_t i) const { return p[i]; }
};
template <class Op, class Src>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(Src src, DevN dn, typename Op::T *partials) {
  using T = typename Op::T;
  const uint64_t n = dn.get();
  __shared__ T lds[kBlock / 64];
  uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  T acc = Op::identity();
#pragma unroll
  for (int i = 0; i < kItems; ++i)
    if (base + i < n) acc = Op::combine(acc, src(base + i));
  T ex, tot;
  block_scan<Op>(acc, &ex, &tot, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Exclusive scan of up to any number of partials with a single block (carry loop).
// Block b scans the b-th array of n partials (partials + b * n).
template <class Op>
__global__ void __launch_bounds__(kBlock) k_scan_partials(typename Op::T *partials, uint64_t n) {
  using T = typename Op::T;
  partials += (uint64_t)blockIdx.x * n;
  __shared__ T lds[kBlock / 64];
  T carry = Op::identity();
  for (uint64_t base = 0; base < n; base += kTile) {
    T items[kItems];
    uint64_t b = base + (uint64_t)threadIdx.x * kItems;
    T acc = Op::identity();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      items[i] = (b + i < n) ? partials[b + i] : Op::identity();
      acc = Op::combine(acc, items[i]);
    }
    T ex, tot;
    block_scan<Op>(acc, &ex, &tot, lds);
    T run = Op::combine(carry, ex);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if (b + i < n) partials[b + i] = run;  // exclusive
      run = Op::combine(run, items[i]);
    }
    carry = Op::combine(carry, tot);
    __syncthreads();
  }
}

// Final pass: exclusive u32 output (exact while the total fits 32 bits); the total
// is written saturated at 0xFFFFFFFF, which callers treat as overflow.
__global__ void __launch_bounds__(kBlock) k_scan_final_u32(const uint32_t *in, uint32_t *out, DevN dn,
                                                            const uint64_t *partials, uint32_t *d_total,
                                                            uint32_t nblocks) {
  __shared__ uint64_t lds[kBlock / 64];
  const uint64_t n = dn.get();
  uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  uint32_t items[kItems];
  uint64_t acc = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    items[i] = (base + i < n) ? in[base + i] : 0u;
    acc += items[i];
  }
  uint64_t ex, tot;
  block_scan<OpAdd64>(acc, &ex, &tot, lds);
  uint64_t run = partials[blockIdx.x] + ex;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (base + i < n) out[base + i] = (uint32_t)run;
    run += items[i];
  }
  if (d_total && blockIdx.x == nblocks - 1 && threadIdx.x == kBlock - 1)
    *d_total = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
}

__global__ void k_fill_words(uint32_t *__restrict__ a, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

void fill_u32(Ctx *ctx, uint32_t *a, uint64_t n, uint32_t v) {
  if (n) launch(ctx, "fill_u32", k_fill_words, dim3((uint32_t)std::min<uint64_t>(div_up(n, 256), 4096)), dim3(256), 0, a, n, v);
}

void exclusive_scan_u32(Ctx *ctx, const uint32_t *in, uint32_t *out, DevN n, uint32_t *d_total, Arena *arena) {
  if (n.cap == 0) {
    if (d_total) fill_u32(ctx, d_total, 1, 0u);
    return;
  }
  uint32_t nb = div_up(n.cap, kTile);
  Buf<uint64_t> partials(ctx, arena, nb);
  launch(ctx, "scan_reduce_u32", k_scan_reduce<OpAdd64, U32Source>, dim3(nb), dim3(kBlock), 0, U32Source{in}, n,
         partials.p);
  launch(ctx, "scan_partials_u32", k_scan_partials<OpAdd64>, dim3(1), dim3(kBlock), 0, partials.p, (uint64_t)nb);
  launch(ctx, "scan_final_u32", k_scan_final_u32, dim3(nb), dim3(kBlock), 0, in, out, n, (const uint64_t *)partials.p,
         d_total, nb);
}

// Both boundary-level scans at once.  Tile b covers positions [b * kTile, min((b + 1) *
// kTile, n)); the left scan walks a tile from its start, the right scan from its end,
// so the right scan's tile states are stored from the last tile down (pr[nb - 1 - b])
// and scanned in that order.
__device__ __forceinline__ LevelState level_item(uint32_t x) { return LevelState{x, 1u << x}; }

__global__ void __launch_bounds__(kBlock) k_level_reduce(const uint8_t *__restrict__ v, uint64_t n,
                                                          LevelState *__restrict__ pl, LevelState *__restrict__ pr,
                                                          uint32_t nb) {
  __shared__ LevelState lds[kBlock / 64];
  const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
  const uint64_t len = n - t0 < (uint64_t)kTile ? n - t0 : (uint64_t)kTile;
  LevelState fa = OpLevel::identity(), ra = OpLevel::identity();
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t j = (uint64_t)threadIdx.x * kItems + i;
    if (j < len) {
      fa = OpLevel::combine(fa, level_item(v[t0 + j]));
      ra = OpLevel::combine(ra, level_item(v[t0 + len - 1 - j]));
    }
  }
  LevelState ex, tot;
  block_scan<OpLevel>(fa, &ex, &tot, lds);
  if (threadIdx.x == 0) pl[blockIdx.x] = tot;
  block_scan<OpLevel>(ra, &ex, &tot, lds);
  if (threadIdx.x == 0) pr[nb - 1 - blockIdx.x] = tot;
}

// Masks from the exclusive tile prefixes.  With `need` (n bytes and more: the split
// flags), left[x] is written only where need[x + 1] is set and right[x] only where
// need[x - 1] is: the positions beside a split boundary, the only ones the node
// emission reads.
__global__ void __launch_bounds__(kBlock) k_level_final(const uint8_t *__restrict__ v, uint64_t n,
                                                         const LevelState *__restrict__ pl,
                                                         const LevelState *__restrict__ pr, uint32_t nb,
                                                         const uint8_t *__restrict__ need, uint32_t *__restrict__ left,
                                                         uint32_t *__restrict__ right) {
  __shared__ LevelState lds[kBlock / 64];
  const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
  const uint64_t len = n - t0 < (uint64_t)kTile ? n - t0 : (uint64_t)kTile;
  const uint64_t j0 = (uint64_t)threadIdx.x * kItems;
  LevelState items[kItems];
  {
    LevelState acc = OpLevel::identity();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      items[i] = (j0 + i < len) ? level_item(v[t0 + j0 + i]) : OpLevel::identity();
      acc = OpLevel::combine(acc, items[i]);
    }
    LevelState ex, tot;
    block_scan<OpLevel>(acc, &ex, &tot, lds);
    LevelState run = OpLevel::combine(pl[blockIdx.x], ex);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      run = OpLevel::combine(run, items[i]);  // inclusive
      const uint64_t x = t0 + j0 + i;
      if (j0 + i < len && (!need || (x + 1 < n && need[x + 1]))) left[x] = run.mask;
    }
  }
  {
    LevelState acc = OpLevel::identity();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      items[i] = (j0 + i < len) ? level_item(v[t0 + len - 1 - j0 - i]) : OpLevel::identity();
      acc = OpLevel::combine(acc, items[i]);
    }
    LevelState ex, tot;
    block_scan<OpLevel>(acc, &ex, &tot, lds);
    LevelState run = OpLevel::combine(pr[nb - 1 - blockIdx.x], ex);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      run = OpLevel::combine(run, items[i]);
      const uint64_t x = t0 + len - 1 - j0 - i;
      if (j0 + i < len && (!need || (x > 0 && need[x - 1]))) right[x] = run.mask;
    }
  }
}

void level_scans(Ctx *ctx, const uint8_t *v, uint64_t n, uint32_t *left_mask, uint32_t *right_mask,
                 const uint8_t *need, Arena *arena) {
  if (n == 0) return;
  const uint32_t nb = div_up(n, kTile);
  Buf<LevelState> partials(ctx, arena, 2 * (uint64_t)nb);
  LevelState *pl = partials.p, *pr = partials.p + nb;
  launch(ctx, "level_reduce", k_level_reduce, dim3(nb), dim3(kBlock), 0, v, n, pl, pr, nb);
  launch(ctx, "scan_partials_level", k_scan_partials<OpLevel>, dim3(2), dim3(kBlock), 0, partials.p, (uint64_t)nb);
  launch(ctx, "level_final", k_level_final, dim3(nb), dim3(kBlock), 0, v, n, (const LevelState *)pl,
         (const LevelState *)pr, nb, need, left_mask, right_mask);
}

// ---------------------------------------------------------------------------
// Radix sort: LSD, 8-bit digits, one launch per digit ("onesweep").
//
// One histogram launch counts every digit position at once; a one-block launch
// turns the counts into each digit's global start.  Each scatter launch then
// ranks a tile of keys in LDS (wavefront ballots), publishes the tile's 256
// digit counts and finds the counts of all earlier tiles by decoupled look-back:
// a tile's status word per digit holds 30 bits of count and a flag, AGG (this
// tile's own count) or INC (count of this tile and every earlier one).  Tile ids
// come from an atomic counter, so every tile a block waits on belongs to a block
// that started earlier and publishes its AGG without waiting on anything.  The
// wait is bounded all the same: past kOsSpinLimit polls the block records an
// error (the host raises it) and carries on.  Status words live in two buffers
// used by alternate passes; a pass clears the other buffer's words of its tile
// for the pass after it (the first pass's buffer is cleared by the histogram
// launch's predecessor).
// ---------------------------------------------------------------------------
constexpr uint32_t kOsAgg = 1u << 30, kOsInc = 2u << 30, kOsCount = (1u << 30) - 1u;
constexpr uint32_t kOsMaxPasses = 8;
constexpr uint32_t kOsSpinLimit = 1u << 22;
constexpr uint32_t kOsLook = 8;  // predecessor status words read per look-back step

struct OsState {
  uint32_t base[kOsMaxPasses][256];  // digit counts, then each digit's global start
  uint32_t tile_counter[kOsMaxPasses];
};

__global__ void __launch_bounds__(kBlock) k_os_init(OsState *__restrict__ st, uint32_t *__restrict__ status0,
                                                     uint64_t nstatus) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t *w = reinterpret_cast<uint32_t *>(st);
  for (uint64_t i = t; i < sizeof(OsState) / 4; i += stride) w[i] = 0u;
  for (uint64_t i = t; i < nstatus; i += stride) status0[i] = 0u;
}

__global__ void __launch_bounds__(kBlock) k_os_hist(const uint64_t *__restrict__ keys, DevN dn, uint32_t passes,
                                                     OsState *__restrict__ st) {
  __shared__ uint32_t h[kOsMaxPasses][256];
  const uint64_t n = dn.get();
  for (uint32_t p = 0; p < passes; ++p) h[p][threadIdx.x] = 0u;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t k = keys[i];
    for (uint32_t p = 0; p < passes; ++p) atomicAdd(&h[p][(uint32_t)(k >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (uint32_t p = 0; p < passes; ++p)
    if (h[p][threadIdx.x]) atomicAdd(&st->base[p][threadIdx.x], h[p][threadIdx.x]);
}

// digit counts -> exclusive starts, one pass per wavefront-sized step of a block scan
__global__ void __launch_bounds__(kBlock) k_os_prefix(OsState *__restrict__ st, uint32_t passes) {
  __shared__ uint64_t lds[kBlock / 64];
  for (uint32_t p = 0; p < passes; ++p) {
    const uint64_t c = st->base[p][threadIdx.x];
    uint64_t ex, tot;
    block_scan<OpAdd64>(c, &ex, &tot, lds);
    st->base[p][threadIdx.x] = (uint32_t)ex;
    __syncthreads();
  }
}

// Scatter tile: 1024 threads (16 wavefronts) x 10 keys = 10,240 keys staged in 138 KiB
// of LDS (one block per CU), so a digit's run in the output averages 40 keys (2.5
// 128-byte lines of keys).  Measured on the batched forest build (39.2M keys, five
// passes): 512 x 8 keys 6.07 ms, 1024 x 8 5.97 ms, 1024 x 10 5.88 ms, 512 x 12 6.68 ms
// (LDS for one block only); writing straight from registers without the LDS stage
// was slower (7.4 ms).
constexpr int kOsBlock = 1024;
constexpr int kOsWaves = kOsBlock / 64;
constexpr int kOsItems = 10;
constexpr int kOsTile = kOsBlock * kOsItems;

// Sort within sets (radix_sort_sets): the keys of set s sit in [off[s], off[s+1]) and
// are sorted there on their low bits only.  Tiles never cross a set boundary (a set's
// last tile may be short); tile_base[s] is the first tile of set s (tile_base[nsets]
// the tile count), tile_set[t] the set of tile t, and base[(s * passes + p) * 256 + d]
// the position where digit d of set s starts in pass p.  A set's first tile publishes
// its counts as inclusive, so no look-back crosses into another set.
struct OsSegments {
  const uint64_t *off;
  const uint32_t *tile_base;
  const uint32_t *tile_set;
  const uint32_t *base;
  const uint32_t *ntiles;  // tile_base + nsets
  uint32_t passes;
  // packed words (ibits > 0): sort bits above the low ibits, which hold the index inside
  // the set; the last pass writes (set << set_shift | word >> ibits, set start + index)
  uint32_t ibits;
  uint32_t set_shift;  // 64: no set bits
};

__global__ void __launch_bounds__(kBlock) k_seg_tile_counts(const uint64_t *__restrict__ off, uint32_t nsets,
                                                            uint32_t *__restrict__ tiles) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s < nsets) tiles[s] = (uint32_t)((off[s + 1] - off[s] + kOsTile - 1) / kOsTile);
}

// one block per set: tile_set over the set's tiles
__global__ void __launch_bounds__(kBlock) k_seg_tile_sets(const uint32_t *__restrict__ tile_base,
                                                          uint32_t *__restrict__ tile_set) {
  const uint32_t s = blockIdx.x;
  for (uint32_t t = tile_base[s] + threadIdx.x; t < tile_base[s + 1]; t += kBlock) tile_set[t] = s;
}

// Digit counts per (set, pass): a block takes kSegHistTiles consecutive tiles, counts
// in LDS and adds its counts to the set's when the set changes and at the end.
constexpr uint32_t kSegHistTiles = 4;
constexpr uint64_t kSegMaxTable = 64ull << 20;  // digit-start table bytes (nsets x passes x 256 words)
__global__ void __launch_bounds__(kBlock) k_seg_hist(const uint64_t *__restrict__ keys, OsSegments sg,
                                                     const uint32_t *__restrict__ ntiles, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[kOsMaxPasses][256];
  const uint32_t nt = *ntiles;
  const uint32_t t0 = blockIdx.x * kSegHistTiles;
  if (t0 >= nt) return;
  const uint32_t t1 = min(t0 + kSegHistTiles, nt);
  for (uint32_t p = 0; p < sg.passes; ++p) h[p][threadIdx.x] = 0u;
  __syncthreads();
  uint32_t cur = sg.tile_set[t0];
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t s = sg.tile_set[t];
    if (s != cur) {  // block-uniform
      __syncthreads();  // every count of the previous set is in
      for (uint32_t p = 0; p < sg.passes; ++p) {
        const uint32_t c = h[p][threadIdx.x];
        if (c) atomicAdd(&hist[((uint64_t)cur * sg.passes + p) * 256 + threadIdx.x], c);
        h[p][threadIdx.x] = 0u;
      }
      cur = s;
      __syncthreads();
    }
    const uint64_t b = sg.off[s] + (uint64_t)(t - sg.tile_base[s]) * kOsTile;
    const uint64_t e = min(b + (uint64_t)kOsTile, sg.off[s + 1]);
    for (uint64_t i = b + threadIdx.x; i < e; i += kBlock) {
      const uint64_t k = keys[i];
      for (uint32_t p = 0; p < sg.passes; ++p) atomicAdd(&h[p][(uint32_t)(k >> (sg.ibits + 8 * p)) & 255u], 1u);
    }
  }
  __syncthreads();
  for (uint32_t p = 0; p < sg.passes; ++p) {
    const uint32_t c = h[p][threadIdx.x];
    if (c) atomicAdd(&hist[((uint64_t)cur * sg.passes + p) * 256 + threadIdx.x], c);
  }
}

// one block per set: counts -> the set's digit starts (in place)
__global__ void __launch_bounds__(kBlock) k_seg_prefix(const uint64_t *__restrict__ off, uint32_t passes,
                                                       uint32_t *__restrict__ hist) {
  __shared__ uint64_t lds[kBlock / 64];
  const uint32_t s = blockIdx.x;
  for (uint32_t p = 0; p < passes; ++p) {
    uint32_t *row = hist + ((uint64_t)s * passes + p) * 256;
    uint64_t ex, tot;
    block_scan<OpAdd64>((uint64_t)row[threadIdx.x], &ex, &tot, lds);
    row[threadIdx.x] = (uint32_t)(off[s] + ex);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kOsBlock) k_os_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                          uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                          DevN dn, uint32_t pass, OsState *__restrict__ st,
                                                          uint32_t *__restrict__ status, uint32_t *__restrict__ status_next,
                                                          uint32_t *__restrict__ err, OsSegments sg) {
  __shared__ uint32_t wcnt[kOsWaves][256];
  __shared__ uint32_t base_off[256];
  __shared__ uint32_t tstart[256];
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t skey[kOsTile];
  __shared__ uint32_t sval[kOsTile];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t shift = 8 * pass + sg.ibits;
  const bool packed = sg.ibits != 0;  // (key, index) words, no values until the last pass
  const bool unpack = packed && pass + 1 == sg.passes;
  const bool digit_thread = threadIdx.x < 256;  // threads 0..255 own one digit each
  if (threadIdx.x == 0) s_tile = atomicAdd(&st->tile_counter[pass], 1u);
  if (digit_thread)
    for (int w = 0; w < kOsWaves; ++w) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t tile_id = s_tile;
  // this tile's keys [tile0, n), whether it starts its sort (no look-back), and where
  // its digits start
  uint64_t tile0, n;
  bool first;
  const uint32_t *digit_base;
  uint32_t seg = 0;
  if (sg.base) {
    if (tile_id >= *sg.ntiles) return;  // block-uniform
    const uint32_t s = sg.tile_set[tile_id];
    seg = s;
    const uint32_t local = tile_id - sg.tile_base[s];
    tile0 = sg.off[s] + (uint64_t)local * kOsTile;
    n = min(tile0 + (uint64_t)kOsTile, sg.off[s + 1]);
    first = local == 0;
    digit_base = sg.base + ((uint64_t)s * sg.passes + pass) * 256;
  } else {
    n = dn.get();
    tile0 = (uint64_t)tile_id * kOsTile;
    first = tile_id == 0;
    digit_base = st->base[pass];
  }
  // tiles past a device-side count have nothing to do; no tile in front of them waits on them
  if (tile0 >= n) return;
  const uint64_t tile = tile0 + (uint64_t)wave * (64 * kOsItems);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint64_t key[kOsItems];
  uint32_t val[kOsItems];
  uint32_t dig[kOsItems];
  uint32_t off[kOsItems];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    uint64_t i = tile + (uint64_t)r * 64 + lane;
    bool valid = i < n;
    key[r] = valid ? kin[i] : 0ull;
    val[r] = (valid && !packed) ? vin[i] : 0u;
    uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
    dig[r] = d;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    uint32_t before = wcnt[wave][d];
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    off[r] = before + rank;
    // the lowest peer publishes the group's count
    if (valid && (peers & lt) == 0ull) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // per digit: waves' offsets inside the tile's run, and the tile's count
  uint32_t run = 0;
  if (digit_thread) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int w = 0; w < kOsWaves; ++w) {
      const uint32_t c = wcnt[w][t];
      wcnt[w][t] = run;
      run += c;
    }
    // look-back: this tile's start of digit d among all tiles
    const uint32_t d = threadIdx.x;
    uint32_t *mine = &status[(uint64_t)tile_id * 256 + d];
    uint32_t excl = 0;
    if (first) {
      __hip_atomic_store(mine, kOsInc | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(mine, kOsAgg | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // each step reads up to kOsLook predecessors at once (independent loads in
      // flight together) and takes the published prefix nearest to this tile
      uint64_t j = tile_id;  // tiles j-1, j-2, ... are still to be taken
      uint32_t spins = 0;
      for (;;) {
        uint32_t sv[kOsLook];
        const uint32_t m = j < kOsLook ? (uint32_t)j : kOsLook;
#pragma unroll
        for (uint32_t q = 0; q < kOsLook; ++q)
          sv[q] = q < m ? __hip_atomic_load(&status[(j - 1 - q) * 256 + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                        : 0u;
        uint32_t used = 0;
        bool done = false;
#pragma unroll
        for (uint32_t q = 0; q < kOsLook; ++q) {
          if (done || q != used || q >= m || (sv[q] & ~kOsCount) == 0u) continue;
          excl += sv[q] & kOsCount;
          used = q + 1;
          done = (sv[q] & kOsInc) != 0u;
        }
        if (done) break;
        j -= used;
        if (used == 0) {  // the nearest remaining tile has not published yet
          if (++spins > kOsSpinLimit) {
            atomicOr(err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __hip_atomic_store(mine, kOsInc | (excl + run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    base_off[d] = digit_base[d] + excl;
    status_next[(uint64_t)tile_id * 256 + d] = 0u;
    // tile-local start of each digit (exclusive scan of the tile counts in digit order)
    uint32_t x = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    tstart[threadIdx.x] = x - run;  // within the wavefront; the waves' sums are added below
  }
  __syncthreads();
  if (digit_thread) {
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += wsum[w];
    tstart[threadIdx.x] += add;
  }
  __syncthreads();
  // stage the tile in digit order in LDS, then write each digit's run with
  // consecutive threads so the global stores coalesce
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const uint64_t i = tile + (uint64_t)r * 64 + lane;
    if (i < n) {
      const uint32_t lp = tstart[dig[r]] + wcnt[wave][dig[r]] + off[r];
      skey[lp] = key[r];
      if (!packed) sval[lp] = val[r];
    }
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)((n - tile0) < (uint64_t)kOsTile ? (n - tile0) : (uint64_t)kOsTile);
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const uint32_t j = (uint32_t)r * kOsBlock + threadIdx.x;
    if (j < cnt) {
      const uint64_t k = skey[j];
      const uint32_t d = (uint32_t)(k >> shift) & 255u;
      const uint32_t pos = base_off[d] + (j - tstart[d]);
      if (!packed) {
        kout[pos] = k;
        vout[pos] = sval[j];
      } else if (!unpack) {
        kout[pos] = k;
      } else {
        const uint64_t set_key = sg.set_shift >= 64 ? 0ull : (uint64_t)seg << sg.set_shift;
        kout[pos] = set_key | (k >> sg.ibits);
        vout[pos] = (uint32_t)(sg.off[seg] + (k & ((1ull << sg.ibits) - 1ull)));
      }
    }
  }
}

void radix_sort_pairs(Ctx *ctx, uint64_t *keys, uint32_t *vals, DevN dn, uint32_t bits, Arena *arena) {
  const uint64_t n = dn.cap;  // buffers and grids: the capacity (the count itself when dn.p is null)
  if (n <= 1 || bits == 0) return;
  if (n > kOsCount) throw Error{OTG_ERR_UNSUPPORTED, "radix sort of 2^30 or more keys"};
  const uint32_t passes = std::min<uint32_t>((bits + 7) / 8, kOsMaxPasses);
  const uint32_t ntiles = div_up(n, kOsTile);
  Buf<uint64_t> k2(ctx, arena, n);
  Buf<uint32_t> v2(ctx, arena, n);
  Buf<OsState> st(ctx, arena, 1);
  Buf<uint32_t> status(ctx, arena, (size_t)2 * 256 * ntiles);
  uint32_t *stat[2] = {status.p, status.p + (size_t)256 * ntiles};
  launch(ctx, "radix_init", k_os_init, dim3(std::min<uint32_t>(div_up((uint64_t)256 * ntiles, kBlock), 1024)),
         dim3(kBlock), 0, st.p, stat[0], (uint64_t)256 * ntiles);
  launch(ctx, "radix_hist", k_os_hist, dim3(std::min<uint32_t>(div_up(n, kBlock), 2048)), dim3(kBlock), 0,
         (const uint64_t *)keys, dn, passes, st.p);
  launch(ctx, "radix_prefix", k_os_prefix, dim3(1), dim3(kBlock), 0, st.p, passes);
  uint64_t *ka = keys, *kb = k2.p;
  uint32_t *va = vals, *vb = v2.p;
  for (uint32_t p = 0; p < passes; ++p) {
    launch(ctx, "radix_scatter", k_os_scatter, dim3(ntiles), dim3(kOsBlock), 0, (const uint64_t *)ka,
           (const uint32_t *)va, kb, vb, dn, p, st.p, stat[p & 1], stat[(p + 1) & 1], ctx->d_err, OsSegments{});
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    OTG_HIP_CHECK(hipMemcpyAsync(keys, ka, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    OTG_HIP_CHECK(hipMemcpyAsync(vals, va, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
}

bool radix_sets_fit(uint64_t n, uint32_t nsets, uint32_t bits) {
  const uint32_t passes = (bits + 7) / 8;
  // every set takes at least one tile: with more sets than the whole-array sort has
  // tiles, most tiles would be nearly empty
  return nsets > 0 && nsets <= div_up(n, kOsTile) && passes <= kOsMaxPasses &&
         (uint64_t)nsets * passes * 256 * 4 <= kSegMaxTable;
}

void radix_sort_sets(Ctx *ctx, uint64_t *ka, uint32_t *va, uint64_t *kb, uint32_t *vb, uint64_t n,
                     const uint64_t *d_off, uint32_t nsets, uint32_t bits, Arena *arena, uint32_t ibits,
                     uint32_t set_shift) {
  if (n == 0 || bits == 0) return;
  if (n > kOsCount) throw Error{OTG_ERR_UNSUPPORTED, "radix sort of 2^30 or more keys"};
  if (!radix_sets_fit(n, nsets, bits)) throw Error{OTG_ERR_INVALID_ARG, "per-set radix tables too large (check radix_sets_fit first)"};
  const uint32_t passes = (bits + 7) / 8;
  const uint64_t ntiles = div_up(n, kOsTile) + nsets;  // bound: one short tile per set
  Buf<OsState> st(ctx, arena, 1);
  Buf<uint32_t> status(ctx, arena, (size_t)2 * 256 * ntiles);
  Buf<uint32_t> tcount(ctx, arena, nsets), tbase(ctx, arena, (uint64_t)nsets + 1), tset(ctx, arena, ntiles);
  Buf<uint32_t> base(ctx, arena, (uint64_t)nsets * passes * 256);
  uint32_t *stat[2] = {status.p, status.p + (size_t)256 * ntiles};
  launch(ctx, "seg_tile_counts", k_seg_tile_counts, dim3(div_up(nsets, kBlock)), dim3(kBlock), 0, d_off, nsets,
         tcount.p);
  exclusive_scan_u32(ctx, tcount.p, tbase.p, nsets, tbase.p + nsets, arena);
  launch(ctx, "seg_tile_sets", k_seg_tile_sets, dim3(nsets), dim3(kBlock), 0, (const uint32_t *)tbase.p, tset.p);
  fill_u32(ctx, base.p, (uint64_t)nsets * passes * 256, 0u);
  if (ibits && (ibits >= 32 || bits + ibits > 64)) throw Error{OTG_ERR_INVALID_ARG, "packed radix words exceed 64 bits"};
  const OsSegments sg{d_off, tbase.p, tset.p, base.p, tbase.p + nsets, passes, ibits, set_shift};
  launch(ctx, "radix_init", k_os_init, dim3(std::min<uint32_t>(div_up((uint64_t)256 * ntiles, kBlock), 1024)),
         dim3(kBlock), 0, st.p, stat[0], (uint64_t)256 * ntiles);
  launch(ctx, "seg_hist", k_seg_hist, dim3(div_up(ntiles, kSegHistTiles)), dim3(kBlock), 0, (const uint64_t *)ka,
         sg, (const uint32_t *)(tbase.p + nsets), base.p);
  launch(ctx, "seg_prefix", k_seg_prefix, dim3(nsets), dim3(kBlock), 0, d_off, passes, base.p);
  for (uint32_t p = 0; p < passes; ++p) {
    launch(ctx, "radix_scatter", k_os_scatter, dim3(ntiles), dim3(kOsBlock), 0, (const uint64_t *)ka,
           (const uint32_t *)va, kb, vb, DevN{n, nullptr}, p, st.p, stat[p & 1], stat[(p + 1) & 1], ctx->d_err, sg);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
}

}  // namespace otg
