// otg_build.hip -- one-launch-per-stage build of OrthoTree's static linear
// tree (core/ot_static_linear_core.h:523-599) for one tree or a forest of many
// independent trees.
//
// The reference builds its LCA-compressed node table with a serial BFS queue
// (Build<true> :417-519).  Here the same table is derived from boundary levels
// of the sorted location codes, with data-parallel passes only:
//
//  1. encode     location key per entity (set | code | depth), id payload
//  2. sort       LDS-ranked LSD radix sort over the key bits in use
//  3. levels     v[i] = 1 + L_i for the boundary between sorted entities i-1, i,
//                where L_i = min(common Morton digits, depth(i-1), depth(i));
//                v = 0 at set boundaries.  The node that boundary i splits is
//                the maximal run around it with interior levels >= L_i, of
//                depth L_i.
//  4. active     boundary i is a split of the final tree iff that run has more
//                than `cap` entities, L_i < maxDepth and (boxes) it is not made
//                of identical keys only (the "all non-refinable" leaf, :465-472).
//                Decided with a bounded window walk over v.
//  5. scans      running-minimum level sets to both sides of every boundary
//                give each node's number of ancestors = its BFS generation.
//  6. nodes      internal node = leftmost active boundary of its run; leaf =
//                segment between consecutive active boundaries, unless it is a
//                single non-refinable box of its parent.  BFS id = rank by
//                (set, generation, position): a radix sort of node keys.
//  7. finalize   entity segments, children ranges (binary search in the sorted
//                node keys), MinPoint geometry (:346-368) / MBR (:280-327).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "otg_device.hpp"
#include "otg_internal.hpp"

namespace otg {

constexpr int kB = 256;

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dmax_std(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dmin_std(double a, double b) { return (b < a) ? b : a; }  // std::min

__device__ __forceinline__ uint32_t set_of(const uint64_t *set_off, uint32_t nsets, uint64_t i) {
  if (nsets == 1) return 0;
  uint32_t lo = 0, hi = nsets;  // find last s with set_off[s] <= i
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (set_off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// set_of for consecutive indices of a wavefront: the first active lane's set is
// found by a wave-uniform search (scalar loads), the others step over the few set
// boundaries inside the wavefront (indices of later lanes are larger).
__device__ __forceinline__ uint32_t set_of_wave(const uint64_t *set_off, uint32_t nsets, uint64_t i) {
  if (nsets == 1) return 0;
  const uint32_t lo32 = __builtin_amdgcn_readfirstlane((uint32_t)i);
  const uint32_t hi32 = __builtin_amdgcn_readfirstlane((uint32_t)(i >> 32));
  uint32_t s = set_of(set_off, nsets, ((uint64_t)hi32 << 32) | lo32);
  while (s + 1 < nsets && set_off[s + 1] <= i) ++s;
  return s;
}

// common leading Morton digits of two codes of a depth-M tree (LCA depth)
template <int D>
__device__ __forceinline__ uint32_t lcp_digits(uint64_t a, uint64_t b, uint32_t M) {
  uint64_t x = a ^ b;
  if (!x) return M;
  return M - (bw64(x) + D - 1) / D;
}

struct KeyView {
  uint32_t dbits, cbits;
  __device__ uint64_t code(uint64_t k) const { return (k >> dbits) & ((cbits >= 64) ? ~0ull : ((1ull << cbits) - 1)); }
  __device__ uint32_t depth(uint64_t k, uint32_t M) const {
    return dbits ? (uint32_t)(k & ((1ull << dbits) - 1)) : M;
  }
  __device__ uint64_t set_base(uint32_t s) const { return (cbits + dbits >= 64) ? 0ull : ((uint64_t)s << (cbits + dbits)); }
  // set id of a sorted key (no search over set offsets needed after the sort)
  __device__ uint32_t set(uint64_t k) const { return (cbits + dbits >= 64) ? 0u : (uint32_t)(k >> (cbits + dbits)); }
};

// ---------------------------------------------------------------------------
// 0. bounding box (internal_geometry_module.h:342-370), exact min/max through
// order-preserving integer images of doubles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ord_bits(double x) {
  uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

template <int D, bool BOX>
__global__ void __launch_bounds__(kB) k_bbox(const double *__restrict__ geom, uint64_t begin, uint64_t end,
                                              unsigned long long *__restrict__ out /* 2D: min ords, max ords */) {
  constexpr int S = BOX ? 2 * D : D;
  uint64_t lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = ~0ull;
    hi[d] = 0ull;
  }
  for (uint64_t i = begin + (uint64_t)blockIdx.x * kB + threadIdx.x; i < end; i += (uint64_t)gridDim.x * kB) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      uint64_t a = ord_bits(geom[i * S + d]);
      uint64_t b = ord_bits(geom[i * S + (BOX ? D + d : d)]);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    for (int off = 32; off > 0; off >>= 1) {
      uint64_t a = __shfl_xor(lo[d], off, 64);
      uint64_t b = __shfl_xor(hi[d], off, 64);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      atomicMin(&out[d], (unsigned long long)lo[d]);
      atomicMax(&out[D + d], (unsigned long long)hi[d]);
    }
  }
}

// ---------------------------------------------------------------------------
// 1. encode
// ---------------------------------------------------------------------------
// error bits
constexpr uint32_t kErrOutOfSpace = 1u;
constexpr uint32_t kErrUnsupported = 2u;

// Items per thread of the encode launch: several points' loads in flight per thread
// (the loads of one point alone leave each wavefront waiting on memory).
template <int S>
constexpr int kPointItems = S <= 3 ? 8 : (S <= 6 ? 4 : 1);

template <int D, bool BOX>
__device__ __forceinline__ void encode_point(const double *p, uint64_t i, uint32_t s,
                                             const SetParams *__restrict__ sets, KeyView kv,
                                             uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                             uint32_t *__restrict__ err, uint32_t ibits, uint64_t s_off) {
  const SetParams &sp = sets[s];
  const uint32_t M = sp.max_depth;
  uint32_t g[D];
  uint64_t code = 0;
  uint32_t depth = M;
  if (!BOX) {
    // si_grid.h:145-175 GetPointGridID
    const uint32_t max_raster_id = (uint32_t)((1ull << M) - 1);
    bool out = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double c = p[d] - sp.space_min[d];
      if (c < -kBaseTol || c > sp.size[d] + kBaseTol) out = true;
      uint32_t r = 0;
      if (c > 0) {
        double x = c * sp.raster[d];
        r = (x >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)x;
        r = r < max_raster_id ? r : max_raster_id;
      }
      g[d] = r;
    }
    if (out) {
      atomicOr(err, kErrOutOfSpace);
      code = 0;
      depth = 0;
    } else {
      code = morton_encode<D>(g);
    }
  } else {
    // si_grid.h:221-285 GetLooseBoxGridData<2.0>, si_mortongrid.h:78-97
    const double *b = p;
    bool inside = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double rmin = sp.space_min[d], rmax = sp.space_max[d];
      double bmin = b[d], bmax = b[D + d];
      if (rmin - kBaseTol > bmin || bmin > rmax + kBaseTol || rmin - kBaseTol > bmax || bmax > rmax + kBaseTol)
        inside = false;
    }
    if (!inside) {
      code = 0;
      depth = 0;
    } else {
      double max_rel = 0.0;
      bool bad = false;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double center = (b[d] + b[D + d]) * 0.5;
        double x = (center - sp.space_min[d]) * sp.raster[d];
        uint32_t gc = (x <= 0.0) ? 0u : ((x >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)x);
        if (gc >> M) bad = true;
        g[d] = gc;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double sz = b[D + d] - b[d];
        max_rel = dmax_std(max_rel, sz * sp.raster[d]);
      }
      double cx = ceil(max_rel);
      uint32_t mrg = (cx >= 4294967295.0) ? 0xFFFFFFFFu : (uint32_t)cx;
      uint32_t level = mrg ? bw64((uint64_t)(mrg - 1)) : 0u;
      if (level > M) bad = true;
      if (bad) {
        atomicOr(err, kErrUnsupported);
        level = 0;
      }
      code = morton_encode<D>(g);
      uint32_t sh = level * D;
      code = (code >> sh) << sh;
      depth = M - level;
    }
  }
  const uint64_t low = (code << kv.dbits) | (BOX ? (uint64_t)depth : 0ull);
  if (ibits) {
    keys[i] = (low << ibits) | (i - s_off);  // packed for the per-set sort
  } else {
    keys[i] = kv.set_base(s) | low;
    vals[i] = (uint32_t)i;
  }
}

// One block encodes kB * kPointItems consecutive points (item r of thread t: point
// base + r * kB + t): every item's coordinates are loaded first, then encoded; the set
// of the block's first point is searched once and stepped forward per item.
template <int D, bool BOX>
__global__ void __launch_bounds__(kB) k_encode(const double *__restrict__ geom, uint64_t n,
                                                const uint64_t *__restrict__ set_off, uint32_t nsets,
                                                const SetParams *__restrict__ sets, KeyView kv,
                                                uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                uint32_t *__restrict__ err, uint32_t ibits) {
  constexpr int S = BOX ? 2 * D : D;
  constexpr int E = kPointItems<S>;
  const uint64_t base = (uint64_t)blockIdx.x * (kB * E);
  if (base >= n) return;
  double c[E][S];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const uint64_t i = base + (uint64_t)r * kB + threadIdx.x;
    if (i < n)
#pragma unroll
      for (int q = 0; q < S; ++q) c[r][q] = geom[i * S + q];
  }
  uint32_t s = set_of_wave(set_off, nsets, base);
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const uint64_t i = base + (uint64_t)r * kB + threadIdx.x;
    if (i >= n) break;
    while (s + 1 < nsets && set_off[s + 1] <= i) ++s;
    encode_point<D, BOX>(c[r], i, s, sets, kv, keys, vals, err, ibits, set_off[s]);
  }
}

// ---------------------------------------------------------------------------
// 3. boundary levels
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kB) k_levels(const uint64_t *__restrict__ keys, uint64_t n,
                                                const uint64_t *__restrict__ set_off, uint32_t nsets,
                                                SetParams *__restrict__ sets, KeyView kv, uint8_t *__restrict__ v) {
  uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (i > n) return;
  uint32_t val = 0;
  uint32_t s = 0;
  bool interior = false;
  if (i > 0 && i < n) {
    const uint64_t ka = keys[i - 1], kb = keys[i];
    s = kv.set(kb);
    if (kv.set(ka) == s) {
      interior = true;
      const uint32_t M = sets[s].max_depth;
      uint32_t L = lcp_digits<D>(kv.code(ka), kv.code(kb), M);
      uint32_t da = kv.depth(ka, M), db = kv.depth(kb, M);
      L = L < da ? L : da;
      L = L < db ? L : db;
      val = L + 1;
    }
  }
  v[i] = (uint8_t)val;
  // a level-0 boundary means the root itself branches (no extra root above a
  // single deeper child); min_lcp becomes 1 then, and stays 255 otherwise
  if (interior && val == 1u) sets[s].min_lcp = 1u;
}

// per-set: does the root split?  (count > cap, and for boxes not every entity at
// depth 0 with one key: ot_static_linear_core.h:441-472)
template <bool BOX>
__global__ void k_root_flags(const uint64_t *__restrict__ keys, SetParams *__restrict__ sets, uint32_t nsets, KeyView kv) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsets) return;
  SetParams &sp = sets[s];
  uint32_t splits = 0;
  if (sp.ent_cnt > sp.cap) {
    splits = 1;
    if (BOX) {
      uint64_t kf = keys[sp.ent_off], kl = keys[sp.ent_off + sp.ent_cnt - 1];
      if (kf == kl && kv.depth(kf, sp.max_depth) == 0) splits = 0;
    }
  }
  sp.root_splits = splits;
}

// ---------------------------------------------------------------------------
// 4. active boundaries
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t lower_bound_u64(const uint64_t *a, uint64_t lo, uint64_t hi, uint64_t x) {
  while (lo < hi) {
    uint64_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint64_t upper_bound_u64(const uint64_t *a, uint64_t lo, uint64_t hi, uint64_t x) {
  while (lo < hi) {
    uint64_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// lower_bound_u64 by galloping: probes 1, 2, 4, ... positions from the low end (fwd)
// or from the high end (back) of [lo, hi), then a binary search inside the last step.
// The same index as lower_bound_u64, found with reads near the end the answer is
// expected at.
__device__ __forceinline__ uint64_t lower_bound_fwd(const uint64_t *a, uint64_t lo, uint64_t hi, uint64_t x) {
  uint64_t step = 1;
  while (lo < hi) {  // a[..lo) < x
    const uint64_t c = (hi - lo > step) ? lo + step - 1 : hi - 1;
    if (a[c] >= x) return lower_bound_u64(a, lo, c, x);
    lo = c + 1;
    step <<= 1;
  }
  return hi;
}
__device__ __forceinline__ uint64_t lower_bound_back(const uint64_t *a, uint64_t lo, uint64_t hi, uint64_t x) {
  uint64_t step = 1;
  while (hi > lo) {  // a[hi..) >= x
    const uint64_t c = (hi - lo > step) ? hi - step : lo;
    if (a[c] < x) return lower_bound_u64(a, c + 1, hi, x);
    hi = c;
    step <<= 1;
  }
  return lo;
}

// Bytes of x below n, as their high bits (every byte and n below 128): the per-byte
// subtraction (x | 0x80) - n never borrows across bytes, so each byte is exact.
__device__ __forceinline__ uint64_t bytes_below(uint64_t x, uint32_t n) {
  const uint64_t H = 0x8080808080808080ull;
  return ~((x | H) - 0x0101010101010101ull * (uint64_t)n) & H;
}

// Number of consecutive levels v[p], v[p+1], ... that are >= vi (stops at the first
// smaller one), counted 8 bytes per load; may overshoot `limit` by up to 7.  The
// walk ends at a set boundary (level 0) at the latest, and v is padded to 8 bytes.
__device__ __forceinline__ uint32_t run_right(const uint8_t *__restrict__ v, uint64_t p, uint32_t vi,
                                              uint32_t limit) {
  uint32_t c = 0;
  while (c < limit) {
    const uint64_t base = p & ~7ull;
    const uint32_t o = (uint32_t)(p - base);
    const uint64_t m = bytes_below(*reinterpret_cast<const uint64_t *>(v + base), vi) & (~0ull << (8 * o));
    if (m) return c + ((uint32_t)__builtin_ctzll(m) >> 3) - o;
    c += 8 - o;
    p = base + 8;
  }
  return c;
}

// Same, walking down from p.
__device__ __forceinline__ uint32_t run_left(const uint8_t *__restrict__ v, uint64_t p, uint32_t vi, uint32_t limit) {
  uint32_t c = 0;
  while (c < limit) {
    const uint64_t base = p & ~7ull;
    const uint32_t o = (uint32_t)(p - base);
    const uint64_t keep = (o == 7) ? ~0ull : ((1ull << (8 * (o + 1))) - 1ull);
    const uint64_t m = bytes_below(*reinterpret_cast<const uint64_t *>(v + base), vi) & keep;
    if (m) return c + o - ((63u - (uint32_t)__builtin_clzll(m)) >> 3);
    c += o + 1;
    if (base == 0) return c;
    p = base - 1;
  }
  return c;
}

template <bool BOX>
__global__ void __launch_bounds__(kB) k_active(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ v,
                                                uint64_t n, KeyView kv, const SetParams *__restrict__ sets,
                                                uint8_t *__restrict__ active) {
  uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (i > n) return;
  uint8_t a = 0;
  const uint32_t vi = (i > 0 && i < n) ? v[i] : 0u;
  if (vi) {
    const uint32_t s = kv.set(keys[i]);
    const SetParams &sp = sets[s];
    const uint32_t L = vi - 1;
    if (L < sp.max_depth) {
      const uint64_t cap = sp.cap;
      // entities of the run around boundary i: the two beside it plus one per
      // neighbouring boundary with a level >= vi, on either side
      const uint32_t lim = (uint32_t)(cap < 0xFFFFFFF0ull ? cap : 0xFFFFFFF0ull);
      uint64_t cnt = 2 + (uint64_t)run_left(v, i - 1, vi, lim);
      if (cnt <= cap) cnt += run_right(v, i + 1, vi, lim);
      if (cnt > cap) {
        a = 1;
        if (BOX && keys[i - 1] == keys[i]) {
          // run of one repeated key at depth L: a leaf holding only
          // non-refinable entities unless the run extends beyond the key block
          const uint64_t sb = sp.ent_off, se = sp.ent_off + sp.ent_cnt;
          uint64_t lo = lower_bound_back(keys, sb, i, keys[i]);
          uint64_t hi = keys[i] == ~0ull ? se : lower_bound_fwd(keys, i, se, keys[i] + 1);  // upper bound
          if (v[lo] < vi && v[hi] < vi) a = 0;
        }
      }
    }
  }
  active[i] = a;
}

// The same test with the levels of a block's 2048 boundaries and a halo on either
// side staged in LDS, together with the minimum level of every 8-byte word: a walk
// finishes its own word byte by byte, then skips 64 levels per step on the word minima
// until a word holds a smaller level, and ends inside that word.  The set of a
// boundary comes from the set offsets, not from its key.  Every set's cap + 64 must
// fit the halo (the host checks): a walk that has not yet reached its limit stays
// inside the window.
constexpr uint32_t kActTile = 2048;  // boundaries per block, 8 per thread
constexpr uint32_t kActHalo = 1024;  // bytes of levels staged on either side
constexpr uint32_t kActWords = (kActTile + 2 * kActHalo) / 8;

// consecutive levels >= vi from window position lp upwards (at least `limit` when
// the run reaches that far)
__device__ __forceinline__ uint32_t lds_run_right(const uint64_t *sv, const uint64_t *smin, uint32_t lp, uint32_t vi,
                                                  uint32_t limit) {
  const uint32_t o = lp & 7u;
  const uint64_t m = bytes_below(sv[lp >> 3], vi) & (~0ull << (8 * o));
  if (m) return ((uint32_t)__builtin_ctzll(m) >> 3) - o;
  uint32_t c = 8 - o;
  uint32_t w = (lp >> 3) + 1;
  while (c < limit) {
    const uint32_t ow = w & 7u;
    const uint64_t mm = bytes_below(smin[w >> 3], vi) & (~0ull << (8 * ow));
    if (mm) {
      const uint32_t wf = (w & ~7u) + ((uint32_t)__builtin_ctzll(mm) >> 3);
      return c + 8 * (wf - w) + ((uint32_t)__builtin_ctzll(bytes_below(sv[wf], vi)) >> 3);
    }
    c += 8 * (8 - ow);
    w = (w & ~7u) + 8;
  }
  return c;
}

// the same from lp downwards
__device__ __forceinline__ uint32_t lds_run_left(const uint64_t *sv, const uint64_t *smin, uint32_t lp, uint32_t vi,
                                                 uint32_t limit) {
  const uint32_t o = lp & 7u;
  const uint64_t keep = (o == 7) ? ~0ull : ((1ull << (8 * (o + 1))) - 1ull);
  const uint64_t m = bytes_below(sv[lp >> 3], vi) & keep;
  if (m) return o - ((63u - (uint32_t)__builtin_clzll(m)) >> 3);
  uint32_t c = o + 1;
  int32_t w = (int32_t)(lp >> 3) - 1;
  while (c < limit) {  // c < limit keeps w >= 0: every level below the window start is counted
    const uint32_t ow = (uint32_t)w & 7u;
    const uint64_t keepw = (ow == 7) ? ~0ull : ((1ull << (8 * (ow + 1))) - 1ull);
    const uint64_t mm = bytes_below(smin[w >> 3], vi) & keepw;
    if (mm) {
      const int32_t wf = (w & ~7) + (int32_t)((63u - (uint32_t)__builtin_clzll(mm)) >> 3);
      const uint64_t mf = bytes_below(sv[wf], vi);
      return c + 8 * (uint32_t)(w - wf) + 7u - ((63u - (uint32_t)__builtin_clzll(mf)) >> 3);
    }
    c += 8 * (ow + 1);
    w = (w & ~7) - 1;
  }
  return c;
}

template <bool BOX>
__global__ void __launch_bounds__(kB) k_active_lds(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ v,
                                                    uint64_t n, const uint64_t *__restrict__ set_off, uint32_t nsets,
                                                    const SetParams *__restrict__ sets, uint8_t *__restrict__ active) {
  __shared__ uint64_t sv[kActWords];
  __shared__ uint64_t smin[kActWords / 8];
  const uint64_t i0 = (uint64_t)blockIdx.x * kActTile;
  const int64_t w0 = (int64_t)i0 - (int64_t)kActHalo;  // byte position of sv[0]
  // levels v[0..n] (bytes past n read as 0: every walk stops at v[n] = 0 anyway)
  for (uint32_t w = threadIdx.x; w < kActWords; w += kB) {
    const int64_t b = w0 + 8 * (int64_t)w;
    uint64_t x = 0;
    if (b >= 0 && (uint64_t)b <= n) {
      if ((uint64_t)b + 8 <= n + 1) {
        x = *reinterpret_cast<const uint64_t *>(v + b);
      } else {
        for (uint64_t q = (uint64_t)b; q <= n; ++q) x |= (uint64_t)v[q] << (8 * (q - (uint64_t)b));
      }
    }
    sv[w] = x;
    uint32_t mn = 255u;
#pragma unroll
    for (int j = 0; j < 8; ++j) mn = min(mn, (uint32_t)(x >> (8 * j)) & 0xFFu);
    reinterpret_cast<uint8_t *>(smin)[w] = (uint8_t)mn;
  }
  __syncthreads();
  // set of the thread's first boundary (the last set starting at or before it), then
  // forward as the boundaries grow
  uint32_t s = (uint32_t)upper_bound_u64(set_off, 0, nsets, i0 + threadIdx.x);
  s = s ? s - 1 : 0u;
#pragma unroll
  for (uint32_t j = 0; j < kActTile / kB; ++j) {
    const uint64_t i = i0 + (uint64_t)j * kB + threadIdx.x;
    if (i > n) break;
    uint8_t a = 0;
    const uint32_t lp = (uint32_t)((int64_t)i - w0);
    const uint32_t vi = (i > 0 && i < n) ? (uint32_t)((sv[lp >> 3] >> (8 * (lp & 7u))) & 0xFFu) : 0u;
    if (vi) {
      while (s + 1 < nsets && set_off[s + 1] <= i) ++s;
      const SetParams &sp = sets[s];
      if (vi - 1 < sp.max_depth) {
        const uint32_t lim = sp.cap;
        uint32_t cnt = 2 + lds_run_left(sv, smin, lp - 1, vi, lim);
        if (cnt <= lim) cnt += lds_run_right(sv, smin, lp + 1, vi, lim);
        if (cnt > lim) {
          a = 1;
          if (BOX && keys[i - 1] == keys[i]) {
            const uint64_t sb = sp.ent_off, se = sp.ent_off + sp.ent_cnt;
            uint64_t lo = lower_bound_back(keys, sb, i, keys[i]);
            uint64_t hi = keys[i] == ~0ull ? se : lower_bound_fwd(keys, i, se, keys[i] + 1);
            if (v[lo] < vi && v[hi] < vi) a = 0;
          }
        }
      }
    }
    active[i] = a;
  }
}

// ---------------------------------------------------------------------------
// 5/6. node events
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t internal_gen(const uint32_t *lm, const uint32_t *rm, const uint8_t *v, uint64_t p) {
  const uint32_t vp = v[p];
  if (vp <= 1) return 0;  // the root (level 0)
  uint32_t anc = (lm[p - 1] | rm[p + 1]) & ((1u << vp) - 1u) & ~3u;
  return (uint32_t)__popc(anc) + 1u;
}

// Which segments are nodes (leaves): all except single non-refinable boxes.
template <bool BOX>
__global__ void __launch_bounds__(kB) k_seg_nodes(const uint32_t *__restrict__ seg_list, DevN dseg,
                                                   const uint8_t *__restrict__ v, const uint64_t *__restrict__ keys,
                                                   const uint64_t *__restrict__ set_off, uint32_t nsets,
                                                   const SetParams *__restrict__ sets, KeyView kv,
                                                   uint32_t *__restrict__ seg_node) {
  uint32_t k = blockIdx.x * kB + threadIdx.x;
  if (k >= dseg.get()) return;
  uint32_t node = 1;
  if (BOX) {
    uint64_t s = seg_list[k], e = seg_list[k + 1];
    if (e - s == 1) {
      uint32_t st = kv.set(keys[s]);
      uint32_t M = sets[st].max_depth;
      uint32_t vs = (sets[st].ent_off == s) ? 0u : v[s];
      uint32_t ve = v[e];
      uint32_t pv = vs > ve ? vs : ve;
      if (pv >= 1 && kv.depth(keys[s], M) + 1 == pv) node = 0;
    }
  }
  seg_node[k] = node;
}

struct NodeKeyLayout {
  uint32_t cobits;  // coordinate bits
  uint32_t gbits;   // generation bits
  __device__ uint64_t make(uint32_t s, uint32_t gen, uint64_t coord) const {
    return ((uint64_t)s << (gbits + cobits)) | ((uint64_t)gen << cobits) | coord;
  }
};

constexpr uint32_t kTypeRoot = 0u, kTypeInternal = 1u, kTypeLeaf = 2u;

__global__ void __launch_bounds__(kB) k_emit_leaves(const uint32_t *__restrict__ seg_list, DevN dseg,
                                                     const uint32_t *__restrict__ seg_node, const uint32_t *__restrict__ seg_node_idx,
                                                     const uint32_t *__restrict__ nint, uint64_t cap, const uint8_t *__restrict__ v,
                                                     const uint32_t *__restrict__ lm, const uint32_t *__restrict__ rm,
                                                     const uint64_t *__restrict__ keys, KeyView kv,
                                                     const SetParams *__restrict__ sets, NodeKeyLayout nk,
                                                     uint64_t *__restrict__ nkeys, uint32_t *__restrict__ ndesc) {
  uint32_t k = blockIdx.x * kB + threadIdx.x;
  if (k >= dseg.get() || !seg_node[k]) return;
  const uint64_t out = (uint64_t)*nint + seg_node_idx[k];
  if (out >= cap) return;  // beyond the capacity (flagged by the caller)
  uint64_t s0 = seg_list[k], e = seg_list[k + 1];
  const uint32_t s = kv.set(keys[s0]);
  const SetParams &sp = sets[s];
  const bool starts = (sp.ent_off == s0);
  const bool ends = (sp.ent_off + sp.ent_cnt == e);
  uint32_t gen;
  if (starts && ends) {
    gen = sp.root_splits ? 1u : 0u;
  } else {
    uint32_t vs = starts ? 0u : v[s0];
    uint32_t ve = ends ? 0u : v[e];
    uint64_t pb = (vs >= ve) ? s0 : e;
    gen = internal_gen(lm, rm, v, pb) + 1u;
  }
  nkeys[out] = nk.make(s, gen, 2 * (s0 - sp.ent_off));
  ndesc[out] = (kTypeLeaf << 30) | k;
}

// Node events at position p in [0, n): bit 0 a segment starts at p, bit 1 an internal
// node is represented by p (the extra root above a deeper single child at a set's
// start, else the leftmost boundary of its node).  A set starts where the level is 0,
// so the key (for the set) is read only there.  The events are never stored: the
// counting launch and the emitting launch both evaluate them from the levels, the
// split flags and the left level masks.
struct EventCalc {
  const uint8_t *v, *active;
  const uint32_t *lm;
  const uint64_t *keys;
  KeyView kv;
  const SetParams *sets;
  __device__ __forceinline__ uint32_t at(uint64_t p, uint32_t vp, uint32_t ap) const {
    if (vp == 0) {
      const SetParams &sp = sets[kv.set(keys[p])];
      return (sp.root_splits && sp.min_lcp > 1u) ? 3u : 1u;
    }
    if (!ap) return 0u;
    return ((lm[p - 1] >> vp) & 1u) ? 1u : 3u;
  }
  // events of the 8 positions [b, b + 8) (those at or past n are 0); b is a multiple
  // of 8, and v and active hold 8 bytes of padding past n
  __device__ __forceinline__ void eight(uint64_t b, uint64_t n, uint32_t e[8]) const {
    if (b >= n) {
#pragma unroll
      for (int i = 0; i < 8; ++i) e[i] = 0u;
      return;
    }
    const uint64_t vw = *reinterpret_cast<const uint64_t *>(v + b);
    const uint64_t aw = *reinterpret_cast<const uint64_t *>(active + b);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      e[i] = (b + i < n) ? at(b + i, (uint32_t)(vw >> (8 * i)) & 0xFFu, (uint32_t)(aw >> (8 * i)) & 0xFFu) : 0u;
  }
};
// segments in the high 32 bits, internal nodes in the low 32 (each total < 2^31)
__device__ __forceinline__ uint64_t event_count(uint32_t e) { return ((uint64_t)(e & 1u) << 32) | (e >> 1); }

constexpr uint32_t kEvTile = kB * 8;  // positions per block, 8 per thread

// Exclusive block-wide sum of one value per thread; *total = the block's sum.
__device__ __forceinline__ uint64_t event_block_scan(uint64_t x, uint64_t *lds /* kB / 64 */, uint64_t *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t inc = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up((unsigned long long)inc, off, 64);
    if (lane >= off) inc += y;
  }
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kB / 64; ++w) {
    if (w < wave) pre += lds[w];
    tot += lds[w];
  }
  __syncthreads();
  *total = tot;
  return pre + inc - x;
}

__global__ void __launch_bounds__(kB) k_event_count(EventCalc ec, uint64_t n, uint64_t *__restrict__ partials) {
  __shared__ uint64_t lds[kB / 64];
  uint32_t e[8];
  ec.eight((uint64_t)blockIdx.x * kEvTile + (uint64_t)threadIdx.x * 8, n, e);
  uint64_t acc = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) acc += event_count(e[i]);
  uint64_t tot;
  (void)event_block_scan(acc, lds, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// One block: the block sums become exclusive block offsets; totals[0] = segments,
// totals[1] = internal nodes.
__global__ void __launch_bounds__(kB) k_event_offsets(uint64_t *__restrict__ partials, uint32_t nb,
                                                       uint32_t *__restrict__ totals) {
  __shared__ uint64_t lds[kB / 64];
  uint64_t carry = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += kEvTile) {
    const uint32_t b = b0 + threadIdx.x * 8;
    uint64_t x[8], acc = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      x[i] = (b + i < nb) ? partials[b + i] : 0ull;
      acc += x[i];
    }
    uint64_t tot;
    uint64_t run = carry + event_block_scan(acc, lds, &tot);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (b + i < nb) partials[b + i] = run;
      run += x[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) {
    totals[0] = (uint32_t)(carry >> 32);
    totals[1] = (uint32_t)carry;
  }
}

// The segment list (segment k starts at seg_list[k]; seg_list[segments] = n) and the
// internal nodes (keys and descriptors at their event rank); entries beyond a
// capacity are dropped (the caller flags the overflow).
__global__ void __launch_bounds__(kB) k_event_emit(EventCalc ec, uint64_t n, const uint64_t *__restrict__ partials,
                                                    const uint32_t *__restrict__ rm, NodeKeyLayout nk,
                                                    uint32_t *__restrict__ seg_list, DevN nseg,
                                                    uint64_t *__restrict__ nkeys, uint32_t *__restrict__ ndesc,
                                                    uint64_t cap_nodes) {
  __shared__ uint64_t lds[kB / 64];
  const uint64_t b = (uint64_t)blockIdx.x * kEvTile + (uint64_t)threadIdx.x * 8;
  uint32_t e[8];
  ec.eight(b, n, e);
  uint64_t acc = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) acc += event_count(e[i]);
  uint64_t tot;
  uint64_t run = partials[blockIdx.x] + event_block_scan(acc, lds, &tot);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (e[i]) {
      const uint64_t p = b + i;
      const uint64_t seg = run >> 32;
      const uint32_t out = (uint32_t)run;
      if (seg < nseg.cap) seg_list[seg] = (uint32_t)p;
      if ((e[i] >> 1) && out < cap_nodes) {
        const uint32_t s = ec.kv.set(ec.keys[p]);
        const SetParams &sp = ec.sets[s];
        if (ec.v[p] == 0) {
          nkeys[out] = nk.make(s, 0, 0);
          ndesc[out] = (kTypeRoot << 30) | (uint32_t)p;
        } else {
          nkeys[out] = nk.make(s, internal_gen(ec.lm, rm, ec.v, p), 2 * (p - sp.ent_off) - 1);
          ndesc[out] = (kTypeInternal << 30) | (uint32_t)p;
        }
      }
    }
    run += event_count(e[i]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) seg_list[nseg.get()] = (uint32_t)n;
}

// counts[3] = nodes = internal nodes (counts[1]) + leaves (counts[2]).  A count beyond
// its capacity sets *overflow and the node count to 0: the table is then empty (every
// set without nodes, so no walk reads it) instead of partly written.
__global__ void k_node_total(uint32_t *__restrict__ counts, uint64_t cap_seg, uint64_t cap_nodes,
                             uint32_t *__restrict__ overflow) {
  const uint64_t nodes = (uint64_t)counts[1] + counts[2];
  const bool over = counts[0] > cap_seg || nodes > cap_nodes;
  counts[3] = over ? 0u : (uint32_t)nodes;
  if (over && overflow) *overflow = 1u;
}

// *overflow when *count > cap
__global__ void k_check_cap(const uint32_t *__restrict__ count, uint64_t cap, uint32_t *__restrict__ overflow) {
  if (*count > cap) *overflow = 1u;
}

// node offsets per set: first node whose key has set id >= s
__global__ void k_set_node_offsets(const uint64_t *__restrict__ nkeys, DevN dnodes, SetParams *__restrict__ sets,
                                   uint32_t nsets, NodeKeyLayout nk) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsets) return;
  const uint64_t nnodes = dnodes.get();
  uint64_t lo = lower_bound_u64(nkeys, 0, nnodes, nk.make(s, 0, 0));
  uint64_t hi = (s + 1 == nsets) ? nnodes : lower_bound_u64(nkeys, 0, nnodes, nk.make(s + 1, 0, 0));
  sets[s].node_off = lo;
  sets[s].node_cnt = hi - lo;
}

// ---------------------------------------------------------------------------
// 7. finalize
// ---------------------------------------------------------------------------
// InitNodeGeometry (MinPoint, :346-368) of a point-tree node: decode the node key
// (sentinel bit above depth * D code bits), cell centre minus half the node size.
template <int D>
__device__ __forceinline__ void point_node_geometry(uint64_t nkey, uint32_t depth, const SetParams &sp, double *nmin,
                                                    double *nsize) {
  const uint32_t M = sp.max_depth;
  uint32_t g[D];
#pragma unroll
  for (int d = 0; d < D; ++d) g[d] = 0;
  uint32_t shift = 0;
  for (uint32_t lvl = M - depth; lvl < M; ++lvl)
#pragma unroll
    for (int d = 0; d < D; ++d, ++shift) g[d] |= (uint32_t)((nkey >> shift) & 1ull) << lvl;
  const double half_grid = (double)(1ull << (M - depth)) * 0.5;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    double center = ((double)g[d] + half_grid) * sp.deraster[d] + sp.space_min[d];
    nmin[d] = center - sp.depth_size[(depth + 1) * D + d];
    nsize[d] = sp.depth_size[depth * D + d];
  }
}

template <int D, bool BOX>
__global__ void __launch_bounds__(kB) k_finalize(const uint64_t *__restrict__ nkeys, const uint32_t *__restrict__ ndesc,
                                                  DevN dnodes, const uint64_t *__restrict__ keys,
                                                  const uint8_t *__restrict__ v, const uint32_t *__restrict__ seg_list,
                                                  const SetParams *__restrict__ sets, KeyView kv, NodeKeyLayout nk,
                                                  NodeInfo *__restrict__ info, uint32_t *__restrict__ rend,
                                                  uint8_t *__restrict__ ndepth, uint64_t *__restrict__ nodekey,
                                                  double *__restrict__ nmin, double *__restrict__ nsize) {
  uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  const uint64_t nnodes = dnodes.get();
  if (x >= nnodes) return;
  const uint64_t key = nkeys[x];
  const uint32_t s = (uint32_t)(key >> (nk.gbits + nk.cobits));
  const uint32_t gen = (uint32_t)((key >> nk.cobits) & ((1u << nk.gbits) - 1));
  const SetParams &sp = sets[s];
  const uint32_t M = sp.max_depth;
  const uint32_t desc = ndesc[x];
  const uint32_t type = desc >> 30;
  const uint32_t pos = desc & 0x3FFFFFFFu;
  const uint64_t sb = sp.ent_off, se = sp.ent_off + sp.ent_cnt;
  uint64_t a, b;
  uint32_t nr = 0, depth;
  uint64_t pivot_code;
  bool internal = true;
  if (type == kTypeRoot) {
    a = sb;
    b = se;
    depth = 0;
    pivot_code = 0;
  } else if (type == kTypeInternal) {
    const uint64_t p = pos;
    const uint32_t L = (uint32_t)v[p] - 1u;
    const uint32_t r = (M - L) * D;  // bits below the node's digits
    const uint64_t code = kv.code(keys[p]);
    const uint64_t P = (r >= 64) ? 0ull : (code >> r);
    const uint64_t lo_code = (r >= 64) ? 0ull : (P << r);
    const uint64_t hi_code = (P + 1) << r;
    const uint64_t lo_key = kv.set_base(s) | (lo_code << kv.dbits) | (BOX ? (uint64_t)L : 0ull);
    a = lower_bound_back(keys, sb, p, lo_key);  // most nodes are small: search near p
    if (hi_code == (1ull << (D * M))) {
      b = se;  // the prefix is the last one of the set
    } else {
      const uint64_t hi_key = kv.set_base(s) + (hi_code << kv.dbits);
      b = lower_bound_fwd(keys, p, se, hi_key);
    }
    if (BOX && keys[a] == lo_key) nr = (uint32_t)(upper_bound_u64(keys, a, b, lo_key) - a);
    depth = L;
    pivot_code = kv.code(keys[a]);
  } else {
    internal = false;
    a = seg_list[pos];
    b = seg_list[pos + 1];
    const uint64_t ka = keys[a];
    pivot_code = kv.code(ka);
    if (a == sb && b == se && !sp.root_splits) {
      depth = 0;  // the root is a leaf
    } else {
      uint32_t L = lcp_digits<D>(pivot_code, kv.code(keys[b - 1]), M);
      uint32_t da = kv.depth(ka, M);
      depth = L < da ? L : da;
    }
  }
  NodeInfo ni;
  ni.ebegin = (uint32_t)a;
  ni.ecount = internal ? nr : (uint32_t)(b - a);
  if (internal) {
    const uint64_t lo = nk.make(s, gen + 1, 2 * (a - sb));
    const uint64_t hi = nk.make(s, gen + 1, 2 * (b - sb));
    const uint64_t c0 = lower_bound_u64(nkeys, x + 1, nnodes, lo);
    const uint64_t c1 = lower_bound_fwd(nkeys, c0, nnodes, hi);  // at most 2^D children
    ni.cbegin = (uint32_t)c0;
    ni.ccount = (uint32_t)(c1 - c0);
  } else {
    ni.cbegin = 0;
    ni.ccount = 0;
  }
  info[x] = ni;
  rend[x] = (uint32_t)b;
  ndepth[x] = (uint8_t)depth;
  // si_morton.h:216-226 node key of the LCA location (pivot code, depth)
  const uint32_t sh = (M - depth) * D;
  const uint64_t nkey = (1ull << (depth * D)) | ((sh >= 64) ? 0ull : (pivot_code >> sh));
  nodekey[x] = nkey;
#pragma unroll
  for (int d = 0; d < D; ++d) nsize[x * D + d] = sp.depth_size[(BOX ? 0 : depth) * D + d];
  if (!BOX) point_node_geometry<D>(nkey, depth, sp, &nmin[x * D], &nsize[x * D]);
}

// ---------------------------------------------------------------------------
// 7b. reference point tree from its finer index (same sorted points, same capacity)
// ---------------------------------------------------------------------------
// The grids differ by an exact power of two per axis, so a node of depth d < M holds
// the same entities in both trees and splits in one iff it splits in the other (count
// > cap, d below both depths); its children group the same entities by the same
// digit, and a child's reference depth is min(its finer depth, M): the common prefix
// of the coarse codes is that of the fine codes cut at M.  The reference nodes are
// therefore the finer roots and the children of finer nodes shallower than M; a kept
// node of finer depth >= M is a reference leaf of depth M holding its subtree's
// entities.  Kept nodes keep all their ancestors, and both tables are numbered by
// (set, generation, position), so the reference numbering is the finer numbering
// restricted to the kept nodes: an exclusive scan of the keep flags.
__global__ void __launch_bounds__(kB) k_ref_mark(const NodeInfo *__restrict__ finfo, const uint8_t *__restrict__ fdepth,
                                                  DevN nf, const uint64_t *__restrict__ fkeys, KeyView fkv,
                                                  const SetParams *__restrict__ rsets, uint32_t *__restrict__ keep) {
  const uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nf.get()) return;
  const NodeInfo ni = finfo[x];
  if (!ni.ccount) return;
  if (fdepth[x] >= rsets[fkv.set(fkeys[ni.ebegin])].max_depth) return;
  for (uint32_t c = 0; c < ni.ccount; ++c) keep[ni.cbegin + c] = 1u;
}

__global__ void k_ref_roots(const SetParams *__restrict__ fsets, uint32_t nsets, uint32_t *__restrict__ keep) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsets && fsets[s].node_cnt) keep[fsets[s].node_off] = 1u;
}

template <int D>
__global__ void __launch_bounds__(kB) k_ref_emit(const NodeInfo *__restrict__ finfo, const uint32_t *__restrict__ frend,
                                                  const uint8_t *__restrict__ fdepth, const uint64_t *__restrict__ fnkey,
                                                  DevN nf, const uint64_t *__restrict__ fkeys, KeyView fkv,
                                                  const SetParams *__restrict__ rsets, const uint32_t *__restrict__ keep,
                                                  const uint32_t *__restrict__ ridx, uint64_t cap, NodeInfo *__restrict__ info,
                                                  uint32_t *__restrict__ rend, uint8_t *__restrict__ ndepth,
                                                  uint64_t *__restrict__ nodekey, double *__restrict__ nmin,
                                                  double *__restrict__ nsize) {
  const uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nf.get() || !keep[x]) return;
  const uint64_t r = ridx[x];
  if (r >= cap) return;  // beyond the capacity (flagged by the caller)
  const NodeInfo fi = finfo[x];
  const SetParams &sp = rsets[fkv.set(fkeys[fi.ebegin])];
  const uint32_t df = fdepth[x];
  const uint32_t dr = df < sp.max_depth ? df : sp.max_depth;
  const bool internal = fi.ccount && df < sp.max_depth;
  NodeInfo ni;
  ni.ebegin = fi.ebegin;
  ni.ecount = internal ? 0u : frend[x] - fi.ebegin;
  ni.cbegin = internal ? ridx[fi.cbegin] : 0u;
  ni.ccount = internal ? fi.ccount : 0u;
  info[r] = ni;
  rend[r] = frend[x];
  ndepth[r] = (uint8_t)dr;
  const uint64_t key = fnkey[x] >> (D * (df - dr));
  nodekey[r] = key;
  point_node_geometry<D>(key, dr, sp, &nmin[r * D], &nsize[r * D]);
}

__global__ void k_ref_sets(const SetParams *__restrict__ fsets, SetParams *__restrict__ rsets, uint32_t nsets,
                           DevN dnf, const uint32_t *__restrict__ ridx, const uint32_t *__restrict__ total,
                           uint64_t cap, const NodeInfo *__restrict__ info) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsets) return;
  const uint64_t nf = dnf.get();
  const uint64_t b = fsets[s].node_off, e = b + fsets[s].node_cnt;
  const uint64_t rb = (b == nf) ? *total : ridx[b];
  const uint64_t re = (e == nf) ? *total : ridx[e];
  const bool over = re > cap;  // beyond the capacity (flagged by the caller): the set gets no nodes
  rsets[s].node_off = over ? 0 : rb;
  rsets[s].node_cnt = over ? 0 : re - rb;
  rsets[s].root_splits = (!over && re > rb && info[rb].ccount) ? 1u : 0u;
}

// sorted SoA geometry and local ids
template <int D, bool BOX>
__global__ void __launch_bounds__(kB) k_gather(const double *__restrict__ geom, const uint32_t *__restrict__ vals, uint64_t n,
                                                const uint64_t *__restrict__ set_off, uint32_t nsets,
                                                double *__restrict__ sorted, uint32_t *__restrict__ ids) {
  uint64_t j = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (j >= n) return;
  constexpr int S = BOX ? 2 * D : D;
  uint32_t id = vals[j];
#pragma unroll
  for (int c = 0; c < S; ++c) sorted[(uint64_t)c * n + j] = geom[(uint64_t)id * S + c];
  // the sets keep their offsets through the sort: position j is in the set of index j
  ids[j] = id - (uint32_t)set_off[set_of_wave(set_off, nsets, j)];
}

// ---------------------------------------------------------------------------
// MBR, bottom-up with arrival counters (InitializeMinimalNodeGeometry :280-327)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kB) k_parents(const NodeInfo *__restrict__ info, uint64_t nnodes,
                                                 uint32_t *__restrict__ parent, uint32_t *__restrict__ pending) {
  uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nnodes) return;
  NodeInfo ni = info[x];
  pending[x] = ni.ccount;
  for (uint32_t c = 0; c < ni.ccount; ++c) parent[ni.cbegin + c] = (uint32_t)x;
}

template <int D>
__device__ void mbr_of_node(uint64_t x, const NodeInfo *info, const double *sorted, uint64_t n, double *nmin,
                            double *nsize) {
  NodeInfo ni = info[x];
  double mn[D], mx[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    mn[d] = 1.7976931348623157e308;
    mx[d] = -1.7976931348623157e308;
  }
  for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      mn[d] = dmin_std(mn[d], sorted[(uint64_t)d * n + j]);
      mx[d] = dmax_std(mx[d], sorted[(uint64_t)(D + d) * n + j]);
    }
  }
  for (uint32_t c = ni.cbegin; c < ni.cbegin + ni.ccount; ++c) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double cmin = nmin[(uint64_t)c * D + d];
      double csz = nsize[(uint64_t)c * D + d];
      mn[d] = dmin_std(mn[d], cmin);
      mx[d] = dmax_std(mx[d], cmin + csz);
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    nmin[x * D + d] = mn[d];
    nsize[x * D + d] = mx[d] - mn[d];
  }
}

template <int D>
__global__ void __launch_bounds__(kB) k_mbr(const NodeInfo *__restrict__ info, uint64_t nnodes, const uint32_t *__restrict__ parent,
                                             uint32_t *__restrict__ pending, const uint8_t *__restrict__ is_root,
                                             const double *__restrict__ sorted, uint64_t n, double *nmin, double *nsize) {
  uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nnodes) return;
  if (info[x].ccount != 0) return;  // start at leaves
  for (;;) {
    mbr_of_node<D>(x, info, sorted, n, nmin, nsize);
    if (is_root[x]) break;
    __threadfence();
    uint32_t p = parent[x];
    uint32_t before = __hip_atomic_fetch_add(&pending[p], 0xFFFFFFFFu, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (before != 1u) break;
    __threadfence();
    x = p;
  }
}

__global__ void k_mark_roots(const SetParams *__restrict__ sets, uint32_t nsets, uint8_t *__restrict__ is_root) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsets) return;
  if (sets[s].node_cnt) is_root[sets[s].node_off] = 1;
}

// leaves holding more than `cap` entities (they sit at max depth or are all non-refinable)
__global__ void __launch_bounds__(kB) k_oversized(const NodeInfo *__restrict__ info, uint64_t nnodes,
                                                   const SetParams *__restrict__ sets, uint32_t nsets,
                                                   uint32_t *__restrict__ flag) {
  uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nnodes) return;
  NodeInfo ni = info[x];
  if (ni.ccount) return;
  uint32_t lo = 0, hi = nsets;  // last set with node_off <= x
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (sets[mid].node_off <= x) lo = mid;
    else hi = mid;
  }
  if (ni.ecount > sets[lo].cap) atomicOr(flag, 1u);
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
uint32_t max_theoretical_depth(int dim) { return dim <= 2 ? (uint32_t)(31 / dim) : (uint32_t)(63 / dim); }

uint32_t estimate_max_depth(uint64_t n, uint64_t cap, int dim) {
  // detail/utils.h:613-622
  if (n <= cap) return 2;
  uint64_t nleaf = n / cap;
  uint32_t est = (uint32_t)(std::log2((double)nleaf) / (double)dim);
  uint32_t hi = max_theoretical_depth(dim);
  return std::clamp<uint32_t>(est, 2u, hi);
}

static void init_set_space(SetParams &sp, int D, bool box, const double *bmin, const double *bmax) {
  // si_grid.h:65-103 and ot_base.h:121-134 (host FP64, IEEE, no contraction)
  volatile double res = (double)(1ull << sp.max_depth);
  for (int d = 0; d < D; ++d) {
    sp.space_min[d] = bmin[d];
    sp.space_max[d] = bmax[d];
    volatile double sz = bmax[d] - bmin[d];
    sp.size[d] = sz;
    sp.deraster[d] = sz / res;
    sp.raster[d] = (sz == 0) ? 1.0 : (res / sz);
  }
  double loose = box ? 2.0 : 1.0;
  for (uint32_t k = 0; k < sp.max_depth + 2; ++k, loose *= 0.5)
    for (int d = 0; d < D; ++d) sp.depth_size[k * D + d] = sp.size[d] * loose;
}

static double ord_to_double(uint64_t u) {
  uint64_t b = (u >> 63) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u;
  double x;
  std::memcpy(&x, &b, 8);
  return x;
}

// Key layout of a tree's sets: set id above D * (largest depth) code bits (and depth
// bits for boxes).
template <int D, bool BOX>
static void set_layout(Tree *t) {
  uint32_t maxM = 1;
  for (const SetParams &sp : t->sets) maxM = std::max(maxM, sp.max_depth);
  const uint32_t nsets = (uint32_t)t->sets.size();
  KeyLayout &L = t->layout;
  L.dbits = BOX ? (32 - __builtin_clz(maxM)) : 0;
  L.cbits = D * maxM;
  L.sbits = nsets > 1 ? (32 - __builtin_clz(nsets - 1)) : 0;
  L.total = L.sbits + L.cbits + L.dbits;
  if (L.total > 64) throw Error{OTG_ERR_KEY_OVERFLOW, "set id and location bits exceed 64"};
}

// Per-set parameters: depth (given or EstimateMaxDepth), capacity, tree box (given,
// or the entities' bounding box: one synchronisation when any set needs it).
template <int D, bool BOX>
static void prepare_sets(Tree *t, const double *dgeom, const uint64_t *offsets, uint32_t nsets, const double *boxes,
                         const uint32_t *depths, const uint32_t *caps) {
  Ctx *ctx = t->ctx;
  t->sets.assign(nsets, SetParams{});
  std::vector<uint32_t> need_bbox;
  for (uint32_t s = 0; s < nsets; ++s) {
    SetParams &sp = t->sets[s];
    sp.ent_off = offsets[s];
    sp.ent_cnt = offsets[s + 1] - offsets[s];
    sp.cap = caps[s];
    if (sp.cap == 0) throw Error{OTG_ERR_INVALID_ARG, "max elements per node must be > 0"};
    uint32_t M = (depths && depths[s]) ? depths[s] : estimate_max_depth(sp.ent_cnt, sp.cap, D);
    if (M < 1 || M > max_theoretical_depth(D) || M > (uint32_t)kMaxDepth)
      throw Error{OTG_ERR_INVALID_ARG, "max depth out of range"};
    sp.max_depth = M;
    sp.min_lcp = 255;
    if (!boxes) need_bbox.push_back(s);
  }
  if (!need_bbox.empty()) {
    Buf<unsigned long long> bb(ctx, (size_t)need_bbox.size() * 2 * D);
    std::vector<unsigned long long> init(need_bbox.size() * 2 * D);
    for (size_t q = 0; q < need_bbox.size(); ++q)
      for (int d = 0; d < D; ++d) {
        init[q * 2 * D + d] = ~0ull;
        init[q * 2 * D + D + d] = 0ull;
      }
    OTG_HIP_CHECK(hipMemcpyAsync(bb.p, init.data(), init.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    for (size_t q = 0; q < need_bbox.size(); ++q) {
      const SetParams &sp = t->sets[need_bbox[q]];
      if (!sp.ent_cnt) continue;
      uint32_t blocks = std::min<uint64_t>(div_up(sp.ent_cnt, kB), 2048);
      launch(ctx, "bbox", k_bbox<D, BOX>, dim3(blocks), dim3(kB), 0, dgeom, sp.ent_off, sp.ent_off + sp.ent_cnt,
             bb.p + q * 2 * D);
    }
    OTG_HIP_CHECK(hipMemcpyAsync(init.data(), bb.p, init.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    for (size_t q = 0; q < need_bbox.size(); ++q) {
      double bmin[kMaxDim], bmax[kMaxDim];
      for (int d = 0; d < D; ++d) {
        bmin[d] = ord_to_double(init[q * 2 * D + d]);
        bmax[d] = ord_to_double(init[q * 2 * D + D + d]);
        if (t->sets[need_bbox[q]].ent_cnt == 0) {
          bmin[d] = 1.7976931348623157e308;
          bmax[d] = -1.7976931348623157e308;
        }
      }
      init_set_space(t->sets[need_bbox[q]], D, BOX, bmin, bmax);
    }
  }
  if (boxes)
    for (uint32_t s = 0; s < nsets; ++s) init_set_space(t->sets[s], D, BOX, &boxes[(size_t)s * 2 * D], &boxes[(size_t)s * 2 * D + D]);
  set_layout<D, BOX>(t);
}

// Memory of a build: the context allocator, or (a build captured into a graph) the
// plan's arena, whose blocks stay with the plan: nothing is allocated or freed while
// capturing, and the tree's arrays keep their addresses on every replay.
static void *build_alloc(Ctx *ctx, Arena *ar, size_t bytes) {
  return ar ? ar->take(std::max<size_t>(bytes, 1)) : ctx->alloc(bytes);
}

// Counts of one node table.  A plain build reads them back and sizes its buffers
// exactly; a captured build sizes them by capacities fixed when its plan was made,
// keeps the counts on the device and flags one beyond its capacity.
struct TableSizes {
  uint32_t *counts = nullptr;    // device: [0] segments, [1] internal nodes, [2] leaves, [3] nodes
  uint64_t cap_seg = 0, cap_nodes = 0;
  uint32_t *overflow = nullptr;  // device flag
};

static void upload_sets(Tree *t) {
  Ctx *ctx = t->ctx;
  const size_t nsets = t->sets.size();
  if (!t->d_sets) t->d_sets = (SetParams *)ctx->alloc(sizeof(SetParams) * nsets);
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_sets, t->sets.data(), sizeof(SetParams) * nsets, hipMemcpyHostToDevice, ctx->stream));
}

// 1-2. location keys of every entity, sorted (t->d_keys); vals = input positions in
// that order; encode error bits land in *err (device).
template <int D, bool BOX>
static void encode_sort(Tree *t, const double *dgeom, const uint64_t *d_off, uint32_t *vals, uint32_t *err,
                        Arena *ar = nullptr) {
  Ctx *ctx = t->ctx;
  const uint64_t n = t->n;
  const KeyView kv{t->layout.dbits, t->layout.cbits};
  t->d_keys = (uint64_t *)build_alloc(ctx, ar, n * 8);
  const uint32_t nsets = (uint32_t)t->sets.size(), low = t->layout.cbits + t->layout.dbits;
  const dim3 grid(div_up(n, kB * kPointItems<BOX ? 2 * D : D>));
  // the points arrive grouped by set: sorting inside each set skips the set-id digits
  // when that saves a pass (C3-like forests: 39 of 47 bits, five passes instead of six)
  if (nsets > 1 && (low + 7) / 8 < (t->layout.total + 7) / 8 && radix_sets_fit(n, nsets, low)) {
    // encode into the buffer that makes the last pass land in t->d_keys and vals
    const bool odd = ((low + 7) / 8) & 1u;
    Buf<uint64_t> k2(ctx, ar, n);
    Buf<uint32_t> v2(ctx, ar, n);
    uint64_t *ka = odd ? k2.p : t->d_keys, *kb = odd ? t->d_keys : k2.p;
    uint32_t *va = odd ? v2.p : vals, *vb = odd ? vals : v2.p;
    // the index inside the set rides below the key bits when both fit one word: 8 bytes
    // per point and pass instead of 12, values written by the last pass only
    uint64_t max_n = 1;
    for (const SetParams &sp : t->sets) max_n = std::max<uint64_t>(max_n, sp.ent_cnt);
    uint32_t ibits = 1;
    while (ibits < 32 && (max_n - 1) >> ibits) ++ibits;
    if (low + ibits > 64) ibits = 0;
    const uint32_t set_shift = t->layout.cbits + t->layout.dbits;
    launch(ctx, "encode", k_encode<D, BOX>, grid, dim3(kB), 0, dgeom, n, d_off, nsets, (const SetParams *)t->d_sets,
           kv, ka, va, err, ibits);
    radix_sort_sets(ctx, ka, va, kb, vb, n, d_off, nsets, low, ar, ibits, set_shift >= 64 ? 64u : set_shift);
    return;
  }
  launch(ctx, "encode", k_encode<D, BOX>, grid, dim3(kB), 0, dgeom, n, d_off, nsets, (const SetParams *)t->d_sets, kv,
         t->d_keys, vals, err, 0u);
  radix_sort_pairs(ctx, t->d_keys, vals, n, t->layout.total, ar);
}

// Keys of the reference tree from the sorted keys of the finer tree over the same
// points: the grids differ by the exact power of two 2^(F - M) per axis, so the code
// at depth M is the code at depth F shifted right by D * (F - M); the order is kept.
template <int D>
__global__ void __launch_bounds__(kB) k_coarse_keys(const uint64_t *__restrict__ fine, uint64_t n, KeyView fkv,
                                                     const SetParams *__restrict__ fsets, KeyView ckv,
                                                     const SetParams *__restrict__ csets, uint64_t *__restrict__ coarse) {
  const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = fine[i];
  const uint32_t s = fkv.set(k);
  const uint32_t sh = D * (fsets[s].max_depth - csets[s].max_depth);
  coarse[i] = ckv.set_base(s) | ((fkv.code(k) >> sh) << ckv.dbits);
}

// The reference tree's keys from its finer index's (same storage order).
template <int D>
static void coarse_keys_of(const Tree *t, const Tree *aux) {
  Ctx *ctx = t->ctx;
  const uint64_t n = t->n;
  t->d_keys = (uint64_t *)ctx->alloc(std::max<uint64_t>(n, 1) * 8);
  launch(ctx, "coarse_keys", k_coarse_keys<D>, dim3(div_up(n, kB)), dim3(kB), 0, (const uint64_t *)aux->d_keys, n,
         KeyView{aux->layout.dbits, aux->layout.cbits}, (const SetParams *)aux->d_sets,
         KeyView{t->layout.dbits, t->layout.cbits}, (const SetParams *)t->d_sets, t->d_keys);
}

template <int D>
static void lazy_coarse_keys(const Tree *t) {
  coarse_keys_of<D>(t, t->knn_tree);
  t->ctx->sync();  // complete before any stream reads them
}

// 3-7. node table from the sorted keys t->d_keys.  herr: host copy of the encode
// error bits, whose copy was queued before; checked after the first synchronisation
// (a plain build).  ts: the capacities and device counts of a captured build.
template <int D, bool BOX>
static void node_table(Tree *t, const uint64_t *d_off, const uint32_t *herr, Arena *ar = nullptr,
                       const TableSizes *ts = nullptr) {
  Ctx *ctx = t->ctx;
  const uint64_t n = t->n;
  const uint32_t nsets = (uint32_t)t->sets.size();
  const KeyLayout &L = t->layout;
  const KeyView kv{L.dbits, L.cbits};
  uint64_t max_n = 1;
  for (const SetParams &sp : t->sets) max_n = std::max<uint64_t>(max_n, sp.ent_cnt);

  // 3. levels (+ per-set minimum level), root flags
  Buf<uint8_t> v(ctx, ar, n + 9);  // levels v[0..n]; 8 bytes of padding for word-wise scans
  launch(ctx, "levels", k_levels<D>, dim3(div_up(n + 1, kB)), dim3(kB), 0, (const uint64_t *)t->d_keys, n, d_off, nsets,
         t->d_sets, kv, v.p);
  launch(ctx, "root_flags", k_root_flags<BOX>, dim3(div_up(nsets, 64)), dim3(64), 0, (const uint64_t *)t->d_keys,
         t->d_sets, nsets, kv);

  // 4. active boundaries
  Buf<uint8_t> active(ctx, ar, n + 9);  // 8 bytes of padding for word-wise reads
  uint64_t max_cap = 0;
  for (const SetParams &sp : t->sets) max_cap = std::max<uint64_t>(max_cap, sp.cap);
  if (max_cap + 64 < kActHalo)
    launch(ctx, "active", k_active_lds<BOX>, dim3(div_up(n + 1, kActTile)), dim3(kB), 0, (const uint64_t *)t->d_keys,
           (const uint8_t *)v.p, n, d_off, nsets, (const SetParams *)t->d_sets, active.p);
  else
    launch(ctx, "active", k_active<BOX>, dim3(div_up(n + 1, kB)), dim3(kB), 0, (const uint64_t *)t->d_keys,
           (const uint8_t *)v.p, n, kv, (const SetParams *)t->d_sets, active.p);

  // 5. level scans
  Buf<uint32_t> lm(ctx, ar, n + 1), rm(ctx, ar, n + 1);
  level_scans(ctx, v.p, n + 1, lm.p, rm.p, active.p, ar);  // only beside split boundaries: all that is read

  // 6. node events: counted per block, block offsets and totals, then (capacities
  // known) the segment list and the internal nodes emitted in one pass
  const EventCalc ec{v.p, active.p, lm.p, t->d_keys, kv, t->d_sets};
  const uint32_t nbe = div_up(n, kEvTile);
  Buf<uint64_t> epart(ctx, ar, std::max<uint32_t>(nbe, 1));
  Buf<uint32_t> own_counts;
  uint32_t *counts = ts ? ts->counts : nullptr;
  if (!counts) {
    own_counts = Buf<uint32_t>(ctx, ar, 4);
    counts = own_counts.p;
  }
  launch(ctx, "event_count", k_event_count, dim3(nbe), dim3(kB), 0, ec, n, epart.p);
  launch(ctx, "event_offsets", k_event_offsets, dim3(1), dim3(kB), 0, epart.p, nbe, counts);
  uint64_t cap_seg, cap_nodes;
  if (!ts) {
    uint32_t ht[2];
    OTG_HIP_CHECK(hipMemcpyAsync(ht, counts, 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (herr && (*herr & kErrOutOfSpace)) throw Error{OTG_ERR_OUT_OF_SPACE, "point outside the given box"};
    if (herr && (*herr & kErrUnsupported)) throw Error{OTG_ERR_UNSUPPORTED, "box level or grid id beyond the tree depth"};
    cap_seg = ht[0];
    cap_nodes = (uint64_t)ht[0] + ht[1];  // refined below
    t->seg_count = cap_seg;
  } else {
    cap_seg = ts->cap_seg;
    cap_nodes = ts->cap_nodes;
  }
  const DevN dseg{cap_seg, counts};

  NodeKeyLayout nk;
  nk.cobits = 64 - __builtin_clzll(2 * max_n + 1);
  nk.gbits = 5;
  if (nk.cobits + nk.gbits + L.sbits > 64) throw Error{OTG_ERR_KEY_OVERFLOW, "node key overflow"};
  Buf<uint32_t> seg_list(ctx, ar, cap_seg + 1), seg_node(ctx, ar, cap_seg), seg_node_idx(ctx, ar, cap_seg);
  Buf<uint64_t> nkeys(ctx, ar, cap_nodes);
  Buf<uint32_t> ndesc(ctx, ar, cap_nodes);
  launch(ctx, "event_emit", k_event_emit, dim3(nbe), dim3(kB), 0, ec, n, (const uint64_t *)epart.p,
         (const uint32_t *)rm.p, nk, seg_list.p, dseg, nkeys.p, ndesc.p, cap_nodes);
  launch(ctx, "seg_nodes", k_seg_nodes<BOX>, dim3(div_up(cap_seg, kB)), dim3(kB), 0, (const uint32_t *)seg_list.p,
         dseg, (const uint8_t *)v.p, (const uint64_t *)t->d_keys, d_off, nsets, (const SetParams *)t->d_sets, kv,
         seg_node.p);
  exclusive_scan_u32(ctx, seg_node.p, seg_node_idx.p, dseg, counts + 2, ar);
  if (!ts) {
    uint32_t h[3];
    OTG_HIP_CHECK(hipMemcpyAsync(h, counts, 12, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    cap_nodes = (uint64_t)h[1] + h[2];
    t->node_count = cap_nodes;
  }
  launch(ctx, "node_total", k_node_total, dim3(1), dim3(1), 0, counts, cap_seg, cap_nodes,
         ts ? ts->overflow : (uint32_t *)nullptr);
  const DevN dnodes{cap_nodes, counts + 3};
  launch(ctx, "emit_leaves", k_emit_leaves, dim3(div_up(cap_seg, kB)), dim3(kB), 0, (const uint32_t *)seg_list.p, dseg,
         (const uint32_t *)seg_node.p, (const uint32_t *)seg_node_idx.p, (const uint32_t *)(counts + 1), cap_nodes,
         (const uint8_t *)v.p, (const uint32_t *)lm.p, (const uint32_t *)rm.p, (const uint64_t *)t->d_keys, kv,
         (const SetParams *)t->d_sets, nk, nkeys.p, ndesc.p);
  radix_sort_pairs(ctx, nkeys.p, ndesc.p, dnodes, nk.cobits + nk.gbits + L.sbits, ar);
  launch(ctx, "set_node_offsets", k_set_node_offsets, dim3(div_up(nsets, 64)), dim3(64), 0, (const uint64_t *)nkeys.p,
         dnodes, t->d_sets, nsets, nk);

  // 7. finalize
  t->d_info = (NodeInfo *)build_alloc(ctx, ar, cap_nodes * sizeof(NodeInfo));
  t->d_rend = (uint32_t *)build_alloc(ctx, ar, cap_nodes * 4);
  t->d_depth = (uint8_t *)build_alloc(ctx, ar, cap_nodes);
  t->d_nkey = (uint64_t *)build_alloc(ctx, ar, cap_nodes * 8);
  t->d_nmin = (double *)build_alloc(ctx, ar, cap_nodes * D * 8);
  t->d_nsize = (double *)build_alloc(ctx, ar, cap_nodes * D * 8);
  launch(ctx, "finalize", k_finalize<D, BOX>, dim3(div_up(cap_nodes, kB)), dim3(kB), 0, (const uint64_t *)nkeys.p,
         (const uint32_t *)ndesc.p, dnodes, (const uint64_t *)t->d_keys, (const uint8_t *)v.p,
         (const uint32_t *)seg_list.p, (const SetParams *)t->d_sets, kv, nk, t->d_info, t->d_rend, t->d_depth,
         t->d_nkey, t->d_nmin, t->d_nsize);
}

// Node table of a point tree from its finer index (k_ref_* above); every set of both
// trees has the same capacity.
template <int D>
static void reference_from_finer(Tree *t, const Tree *aux, Arena *ar = nullptr, const TableSizes *fts = nullptr,
                                 const TableSizes *rts = nullptr) {
  Ctx *ctx = t->ctx;
  const DevN nf = fts ? DevN{fts->cap_nodes, fts->counts + 3} : DevN{aux->node_count, nullptr};
  const uint32_t nsets = (uint32_t)t->sets.size();
  const KeyView fkv{aux->layout.dbits, aux->layout.cbits};
  Buf<uint32_t> keep(ctx, ar, nf.cap), ridx(ctx, ar, nf.cap), own_total;
  uint32_t *total = rts ? rts->counts + 3 : nullptr;
  if (!total) {
    own_total = Buf<uint32_t>(ctx, ar, 1);
    total = own_total.p;
  }
  fill_u32(ctx, keep.p, nf.cap, 0u);
  launch(ctx, "ref_mark", k_ref_mark, dim3(div_up(nf.cap, kB)), dim3(kB), 0, (const NodeInfo *)aux->d_info,
         (const uint8_t *)aux->d_depth, nf, (const uint64_t *)aux->d_keys, fkv, (const SetParams *)t->d_sets, keep.p);
  launch(ctx, "ref_roots", k_ref_roots, dim3(div_up(nsets, 64)), dim3(64), 0, (const SetParams *)aux->d_sets, nsets,
         keep.p);
  exclusive_scan_u32(ctx, keep.p, ridx.p, nf, total, ar);
  uint64_t cap;
  if (!rts) {
    uint32_t nr = 0;
    OTG_HIP_CHECK(hipMemcpyAsync(&nr, total, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    cap = nr;
    t->node_count = nr;
  } else {
    cap = rts->cap_nodes;
    launch(ctx, "check_cap", k_check_cap, dim3(1), dim3(1), 0, (const uint32_t *)total, cap, rts->overflow);
  }
  t->d_info = (NodeInfo *)build_alloc(ctx, ar, std::max<uint64_t>(cap, 1) * sizeof(NodeInfo));
  t->d_rend = (uint32_t *)build_alloc(ctx, ar, std::max<uint64_t>(cap, 1) * 4);
  t->d_depth = (uint8_t *)build_alloc(ctx, ar, std::max<uint64_t>(cap, 1));
  t->d_nkey = (uint64_t *)build_alloc(ctx, ar, std::max<uint64_t>(cap, 1) * 8);
  t->d_nmin = (double *)build_alloc(ctx, ar, std::max<uint64_t>(cap, 1) * D * 8);
  t->d_nsize = (double *)build_alloc(ctx, ar, std::max<uint64_t>(cap, 1) * D * 8);
  launch(ctx, "ref_emit", k_ref_emit<D>, dim3(div_up(nf.cap, kB)), dim3(kB), 0, (const NodeInfo *)aux->d_info,
         (const uint32_t *)aux->d_rend, (const uint8_t *)aux->d_depth, (const uint64_t *)aux->d_nkey, nf,
         (const uint64_t *)aux->d_keys, fkv, (const SetParams *)t->d_sets, (const uint32_t *)keep.p,
         (const uint32_t *)ridx.p, cap, t->d_info, t->d_rend, t->d_depth, t->d_nkey, t->d_nmin, t->d_nsize);
  launch(ctx, "ref_sets", k_ref_sets, dim3(div_up(nsets, 64)), dim3(64), 0, (const SetParams *)aux->d_sets, t->d_sets,
         nsets, nf, (const uint32_t *)ridx.p, (const uint32_t *)total, cap, (const NodeInfo *)t->d_info);
}

// Sorted SoA geometry and local ids of the entity storage.
// keys_of: the tree whose sorted keys give each position's set (a shared sort's
// finer index, whose order is the storage's; null: t).
template <int D, bool BOX>
static void gather_entities(Tree *t, const double *dgeom, const uint32_t *vals, const uint64_t *d_off,
                            Arena *ar = nullptr) {
  Ctx *ctx = t->ctx;
  constexpr int S = BOX ? 2 * D : D;
  const uint64_t n = t->n;
  t->d_geom = (double *)build_alloc(ctx, ar, n * S * 8);
  t->d_ids = (uint32_t *)build_alloc(ctx, ar, n * 4);
  launch(ctx, "gather", k_gather<D, BOX>, dim3(div_up(n, kB)), dim3(kB), 0, dgeom, vals, n, d_off,
         (uint32_t)t->sets.size(), t->d_geom, t->d_ids);
}

template <int D>
static void build_mbr(Tree *t) {
  Ctx *ctx = t->ctx;
  const uint64_t nnodes = t->node_count;
  const uint32_t nsets = (uint32_t)t->sets.size();
  Buf<uint32_t> parent(ctx, nnodes), pending(ctx, nnodes);
  Buf<uint8_t> is_root(ctx, nnodes);
  OTG_HIP_CHECK(hipMemsetAsync(is_root.p, 0, nnodes, ctx->stream));
  launch(ctx, "parents", k_parents, dim3(div_up(nnodes, kB)), dim3(kB), 0, (const NodeInfo *)t->d_info, nnodes,
         parent.p, pending.p);
  launch(ctx, "mark_roots", k_mark_roots, dim3(div_up(nsets, 64)), dim3(64), 0, (const SetParams *)t->d_sets, nsets,
         is_root.p);
  launch(ctx, "mbr", k_mbr<D>, dim3(div_up(nnodes, kB)), dim3(kB), 0, (const NodeInfo *)t->d_info, nnodes,
         (const uint32_t *)parent.p, pending.p, (const uint8_t *)is_root.p, (const double *)t->d_geom, t->n,
         t->d_nmin, t->d_nsize);
}

// Point sets whose max-depth cells are expected to overflow: n / cap above an eighth
// of the 2^(D M) cells.  EstimateMaxDepth picks M with n / cap >= 2^(D M), so every
// set of estimated depth qualifies; a set that does not is checked after its build.
static bool expects_overfull_leaves(const SetParams &sp, int D) {
  const uint32_t bits = D * sp.max_depth;
  if (bits >= 60) return false;
  return (double)sp.ent_cnt / (double)sp.cap > (double)(1ull << bits) / 8.0;
}

// Deepest finer index the key allows for a tree of these sets.
static uint32_t finer_depth(int D, uint32_t sbits) {
  return std::min<uint32_t>({(uint32_t)kMaxDepth, max_theoretical_depth(D), (64 - sbits) / (uint32_t)D});
}

// Depth of a set's finer index: enough levels below the reference depth M for its
// n / cap leaves plus 12 bits of room for uneven density (C3 sets: M = 4, depth 13;
// a tree as deep as the key allows sorts 62 instead of 47 key bits and answered the
// batched kNN and range queries no faster), at most fmax.
static uint32_t finer_set_depth(const SetParams &sp, int D, uint32_t fmax) {
  const double leaves = std::max(1.0, (double)sp.ent_cnt / (double)sp.cap);
  const uint32_t extra = (uint32_t)std::ceil((std::log2(leaves) + 12.0) / (double)D);
  return std::max(sp.max_depth, std::min(fmax, sp.max_depth + extra));
}

static Tree *new_finer_tree(const Tree *t, uint32_t fmax) {
  Tree *aux = new Tree();
  aux->ctx = t->ctx;
  aux->kind = t->kind;
  aux->dim = t->dim;
  aux->n = t->n;
  aux->sets = t->sets;
  for (SetParams &sp : aux->sets) {
    sp.max_depth = finer_set_depth(sp, t->dim, fmax);
    if (t->ctx->tuning.index_capacity) sp.cap = t->ctx->tuning.index_capacity;
    sp.min_lcp = 255;
    sp.root_splits = 0;
    double bmin[kMaxDim], bmax[kMaxDim];
    for (int d = 0; d < t->dim; ++d) {
      bmin[d] = sp.space_min[d];
      bmax[d] = sp.space_max[d];
    }
    init_set_space(sp, t->dim, false, bmin, bmax);
  }
  return aux;
}

template <int D, bool BOX>
static void build_impl(Tree *t, const double *geom_in, int mem, const uint64_t *offsets, uint32_t nsets,
                       const double *boxes, const uint32_t *depths, const uint32_t *caps, bool allow_aux = true) {
  Ctx *ctx = t->ctx;
  constexpr int S = BOX ? 2 * D : D;
  const uint64_t n = offsets[nsets];
  t->n = n;
  if (n >= (1ull << 30)) throw Error{OTG_ERR_UNSUPPORTED, "more than 2^30 entities"};

  // entity geometry on the device (input order)
  Buf<double> dgeom_own;
  const double *dgeom = geom_in;
  if (mem == OTG_MEM_HOST) {
    dgeom_own = Buf<double>(ctx, n * S);
    if (n) OTG_HIP_CHECK(hipMemcpyAsync(dgeom_own.p, geom_in, n * S * 8, hipMemcpyHostToDevice, ctx->stream));
    dgeom = dgeom_own.p;
  }
  prepare_sets<D, BOX>(t, dgeom, offsets, nsets, boxes, depths, caps);
  Buf<uint64_t> d_off(ctx, nsets + 1);
  OTG_HIP_CHECK(hipMemcpyAsync(d_off.p, offsets, (nsets + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
  upload_sets(t);
  if (n == 0) {
    t->node_count = 0;
    return;
  }

  // points: a finer index for kNN / range when max-depth leaves overflow.  When that
  // is expected, one sort by the finer keys serves both trees (shared storage).
  const uint32_t fmax = finer_depth(D, t->layout.sbits);
  bool finer_possible = false, shared = false;
  if (!BOX && allow_aux)
    for (const SetParams &sp : t->sets) {
      finer_possible |= sp.max_depth < fmax;
      shared |= sp.max_depth < fmax && expects_overfull_leaves(sp, D);
    }

  Buf<uint32_t> err(ctx, 1), vals(ctx, n);
  OTG_HIP_CHECK(hipMemsetAsync(err.p, 0, 4, ctx->stream));
  uint32_t herr = 0;
  Tree *aux = nullptr;
  try {
    if (shared) {
      aux = new_finer_tree(t, fmax);
      set_layout<D, false>(aux);
      upload_sets(aux);
      encode_sort<D, false>(aux, dgeom, d_off.p, vals.p, err.p);
      OTG_HIP_CHECK(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, ctx->stream));
      bool same_caps = true;
      for (uint32_t q = 0; q < nsets; ++q) same_caps &= aux->sets[q].cap == t->sets[q].cap;
      // the reference keys: made now when the reference table needs them, else on first
      // use (export, reference entity order) from the finer keys
      if (!same_caps) coarse_keys_of<D>(t, aux);
      node_table<D, false>(aux, d_off.p, &herr);
      if (same_caps) reference_from_finer<D>(t, aux);
      else node_table<D, false>(t, d_off.p, nullptr);
      gather_entities<D, false>(t, dgeom, vals.p, d_off.p);
      aux->d_geom = t->d_geom;
      aux->d_ids = t->d_ids;
      aux->shares_entities = true;
      t->fine_order = true;
      if (same_caps) t->lazy_keys = &lazy_coarse_keys<D>;
    } else {
      encode_sort<D, BOX>(t, dgeom, d_off.p, vals.p, err.p);
      OTG_HIP_CHECK(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, ctx->stream));
      node_table<D, BOX>(t, d_off.p, &herr);
      gather_entities<D, BOX>(t, dgeom, vals.p, d_off.p);
      if (BOX) build_mbr<D>(t);
    }
    for (Tree *x = t; x; x = (x == t ? aux : nullptr))
      OTG_HIP_CHECK(hipMemcpyAsync(x->sets.data(), x->d_sets, sizeof(SetParams) * nsets, hipMemcpyDeviceToHost,
                                   ctx->stream));
    ctx->sync();
  } catch (...) {
    delete aux;
    throw;
  }
  t->device_bytes = n * (S * 8 + 4 + 8) + t->node_count * (sizeof(NodeInfo) + 4 + 1 + 8 + 2 * D * 8);
  if (aux) {
    aux->device_bytes = n * 8 + aux->node_count * (sizeof(NodeInfo) + 4 + 1 + 8 + 2 * D * 8);
    t->knn_tree = aux;
    t->device_bytes += aux->device_bytes;
    return;
  }

  if (finer_possible) {
    // not expected, but max-depth leaves may still overflow (clustered points)
    Buf<uint32_t> flag(ctx, 1);
    OTG_HIP_CHECK(hipMemsetAsync(flag.p, 0, 4, ctx->stream));
    launch(ctx, "oversized", k_oversized, dim3(div_up(t->node_count, kB)), dim3(kB), 0, (const NodeInfo *)t->d_info,
           t->node_count, (const SetParams *)t->d_sets, nsets, flag.p);
    uint32_t hflag = 0;
    OTG_HIP_CHECK(hipMemcpyAsync(&hflag, flag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (hflag) {
      std::vector<double> sboxes((size_t)nsets * 2 * D);
      std::vector<uint32_t> fdepth(nsets);
      for (uint32_t s = 0; s < nsets; ++s) {
        for (int d = 0; d < D; ++d) {
          sboxes[(size_t)s * 2 * D + d] = t->sets[s].space_min[d];
          sboxes[(size_t)s * 2 * D + D + d] = t->sets[s].space_max[d];
        }
        fdepth[s] = finer_set_depth(t->sets[s], D, fmax);
      }
      Tree *x = new Tree();
      x->ctx = ctx;
      x->kind = t->kind;
      x->dim = t->dim;
      std::vector<uint32_t> icaps(caps, caps + nsets);
      if (ctx->tuning.index_capacity) std::fill(icaps.begin(), icaps.end(), ctx->tuning.index_capacity);
      try {
        build_impl<D, false>(x, dgeom, OTG_MEM_DEVICE, offsets, nsets, sboxes.data(), fdepth.data(), icaps.data(),
                             false);
      } catch (...) {
        delete x;
        throw;
      }
      t->knn_tree = x;
      t->device_bytes += x->device_bytes;
    }
  }
}

// ---------------------------------------------------------------------------
// reference storage order of a shared-sort point tree (equal keys by id)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kB) k_ref_ids(const uint32_t *__restrict__ ids, uint64_t n, uint64_t *__restrict__ keys,
                                                 uint32_t *__restrict__ pos) {
  const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  keys[i] = ids[i];
  pos[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(kB) k_ref_keys(const uint64_t *__restrict__ coarse, const uint32_t *__restrict__ pos,
                                                  uint64_t n, uint64_t *__restrict__ keys) {
  const uint64_t j = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (j < n) keys[j] = coarse[pos[j]];
}

template <int D>
__global__ void __launch_bounds__(kB) k_ref_gather(const double *__restrict__ geom, const uint32_t *__restrict__ ids,
                                                    const uint32_t *__restrict__ pos, uint64_t n,
                                                    double *__restrict__ geom_out, uint32_t *__restrict__ ids_out) {
  const uint64_t j = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (j >= n) return;
  const uint64_t p = pos[j];
#pragma unroll
  for (int c = 0; c < D; ++c) geom_out[(uint64_t)c * n + j] = geom[(uint64_t)c * n + p];
  ids_out[j] = ids[p];
}

// Stable sort of the storage by local id, then stable by location key: equal keys by
// id, the order the reference's Create yields from entities taken in id order.
static void make_reference_order(const Tree *t) {
  Ctx *ctx = t->ctx;
  const uint64_t n = t->n;
  uint64_t max_n = 1;
  for (const SetParams &sp : t->sets) max_n = std::max<uint64_t>(max_n, sp.ent_cnt);
  Buf<uint64_t> keys(ctx, n);
  Buf<uint32_t> pos(ctx, n);
  launch(ctx, "ref_ids", k_ref_ids, dim3(div_up(n, kB)), dim3(kB), 0, (const uint32_t *)t->d_ids, n, keys.p, pos.p);
  radix_sort_pairs(ctx, keys.p, pos.p, n, 64 - __builtin_clzll(max_n));
  launch(ctx, "ref_keys", k_ref_keys, dim3(div_up(n, kB)), dim3(kB), 0, (const uint64_t *)t->d_keys,
         (const uint32_t *)pos.p, n, keys.p);
  radix_sort_pairs(ctx, keys.p, pos.p, n, t->layout.total);
  double *geom = (double *)ctx->alloc(n * t->dim * 8);
  uint32_t *ids = (uint32_t *)ctx->alloc(n * 4);
  switch (t->dim) {
#define OTG_REF(DD)                                                                                                  \
  case DD:                                                                                                           \
    launch(ctx, "ref_gather", k_ref_gather<DD>, dim3(div_up(n, kB)), dim3(kB), 0, (const double *)t->d_geom,          \
           (const uint32_t *)t->d_ids, (const uint32_t *)pos.p, n, geom, ids);                                      \
    break;
    OTG_REF(2)
    OTG_REF(3)
    OTG_REF(4)
    OTG_REF(5)
    OTG_REF(6)
#undef OTG_REF
  }
  TreeView hv = view_of(t);
  hv.geom = geom;
  hv.ids = ids;
  TreeView *dv = (TreeView *)ctx->alloc(sizeof(TreeView));
  OTG_HIP_CHECK(hipMemcpyAsync(dv, &hv, sizeof(TreeView), hipMemcpyHostToDevice, ctx->stream));
  ctx->sync();
  t->d_geom_ref = geom;
  t->d_ids_ref = ids;
  t->d_view_ref = dv;
}

void ensure_keys(const Tree *t) {
  if (!t->lazy_keys || t->n == 0) return;
  std::lock_guard<std::mutex> g(t->ref_mu);
  if (!t->d_keys) t->lazy_keys(t);
}

TreeView reference_view_of(const Tree *t) {
  TreeView hv = view_of(t);
  if (!t->fine_order || t->n == 0) return hv;
  ensure_keys(t);
  std::lock_guard<std::mutex> g(t->ref_mu);
  if (!t->d_view_ref) make_reference_order(t);
  hv.geom = t->d_geom_ref;
  hv.ids = t->d_ids_ref;
  return hv;
}

const TreeView *reference_device_view(const Tree *t) {
  if (!t->fine_order || t->n == 0) return t->d_view;
  reference_view_of(t);
  return t->d_view_ref;
}

__global__ void __launch_bounds__(kB) k_nonzero_u32(const uint32_t *__restrict__ in, uint64_t n, uint32_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) out[i] = in[i] ? 1u : 0u;
}

__global__ void __launch_bounds__(kB) k_add_u32(uint32_t *__restrict__ a, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) a[i] += v;
}

// ---------------------------------------------------------------------------
// SEQ-path numbering: Build<ARE_LOCATIONS_SORTED = false> (ot_static_linear_core.h:417-519)
// ---------------------------------------------------------------------------
// The unsorted build splits each node's entity range, in the order the entities
// have reached, by repeated partitions: for boxes first the entities at the node's
// depth (std::partition, :449-459), then one child at a time, the child of the
// range's first entity (the pivot) moved to the front (Partitioning::partition,
// detail/partitioning.h:215-241).  Both are the two-pointer scheme, whose result is
// this: with T the entities that belong in front, the i-th entity from the left of
// [0, T) that does not belong there swaps with the i-th entity from the right of
// [T, n) that does.  A block of 256 threads does one node: it ranks those entities
// with ballots and swaps the pairs.  Each BFS level is one launch; a child is
// recorded at its first position, so the next level, in position order, is the
// reference's queue order (parents in order, each parent's children in pivot order).
constexpr int kSB = 256;

__device__ __forceinline__ uint32_t seq_rank(bool flag, uint32_t *s_w, uint32_t &total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t b = __ballot(flag);
  const uint32_t in_wave = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) s_w[wave] = (uint32_t)__popcll(b);
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kSB / 64; ++w) {
    const uint32_t c = s_w[w];
    if ((uint32_t)w < wave) off += c;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return off + in_wave;
}

struct SeqPred {
  bool by_depth;     // boxes' first partition: entity depth == node depth
  uint32_t depth;
  uint32_t shift;    // child partitions: digit at the examined level equals `flag`
  uint64_t flag;
  uint64_t mask;
  __device__ bool operator()(uint64_t key, KeyView kv, uint32_t M) const {
    return by_depth ? (kv.depth(key, M) == depth) : (((kv.code(key) >> shift) & mask) == flag);
  }
};

struct SeqAcc {
  uint32_t count;
  uint64_t lo, hi;  // smallest and largest code of the front part
  uint32_t mind;    // smallest entity depth of the front part
};

// Partition [lo, hi) of (keys, ids) by pred; returns the front part's count and
// the reduction of its codes.  ff / bt: position scratch over [lo, hi).
__device__ SeqAcc seq_partition(uint64_t *keys, uint32_t *ids, uint32_t lo, uint32_t hi, SeqPred pred, KeyView kv,
                                uint32_t M, uint32_t *ff, uint32_t *bt, uint32_t *s_w, uint64_t *s_r) {
  // front part size and its code range
  uint32_t cnt = 0, mind = 255;
  uint64_t cmin = ~0ull, cmax = 0;
  for (uint32_t p = lo + threadIdx.x; p < hi; p += kSB) {
    const uint64_t k = keys[p];
    if (pred(k, kv, M)) {
      ++cnt;
      const uint64_t c = kv.code(k);
      cmin = c < cmin ? c : cmin;
      cmax = c > cmax ? c : cmax;
      const uint32_t dk = kv.depth(k, M);
      mind = dk < mind ? dk : mind;
    }
  }
  for (int o = 32; o >= 1; o >>= 1) {
    cnt += __shfl_xor(cnt, o, 64);
    const uint64_t a = __shfl_xor(cmin, o, 64), b = __shfl_xor(cmax, o, 64);
    const uint32_t m = __shfl_xor(mind, o, 64);
    cmin = a < cmin ? a : cmin;
    cmax = b > cmax ? b : cmax;
    mind = m < mind ? m : mind;
  }
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0) {
    s_r[4 * wave + 0] = cnt;
    s_r[4 * wave + 1] = cmin;
    s_r[4 * wave + 2] = cmax;
    s_r[4 * wave + 3] = mind;
  }
  __syncthreads();
  SeqAcc acc{0, ~0ull, 0, 255};
#pragma unroll
  for (int w = 0; w < kSB / 64; ++w) {
    acc.count += (uint32_t)s_r[4 * w + 0];
    acc.lo = s_r[4 * w + 1] < acc.lo ? s_r[4 * w + 1] : acc.lo;
    acc.hi = s_r[4 * w + 2] > acc.hi ? s_r[4 * w + 2] : acc.hi;
    acc.mind = (uint32_t)s_r[4 * w + 3] < acc.mind ? (uint32_t)s_r[4 * w + 3] : acc.mind;
  }
  __syncthreads();
  const uint32_t T = acc.count, mid = lo + T;
  // entities of [lo, mid) that leave, left to right
  uint32_t m = 0;
  for (uint32_t p0 = lo; p0 < mid; p0 += kSB) {
    const uint32_t p = p0 + threadIdx.x;
    const bool f = p < mid && !pred(keys[p], kv, M);
    uint32_t tot;
    const uint32_t r = seq_rank(f, s_w, tot);
    if (f) ff[lo + m + r] = p;
    m += tot;
  }
  // entities of [mid, hi) that enter, right to left
  uint32_t m2 = 0;
  for (int64_t q0 = (int64_t)hi - 1; q0 >= (int64_t)mid; q0 -= kSB) {
    const int64_t p = q0 - (int64_t)threadIdx.x;
    const bool f = p >= (int64_t)mid && pred(keys[p], kv, M);
    uint32_t tot;
    const uint32_t r = seq_rank(f, s_w, tot);
    if (f) bt[lo + m2 + r] = (uint32_t)p;
    m2 += tot;
  }
  __syncthreads();  // every pair position written before any is read
  // m == m2: as many leave the front as enter it
  for (uint32_t r = threadIdx.x; r < m; r += kSB) {
    const uint32_t i = ff[lo + r], j = bt[lo + r];
    const uint64_t ki = keys[i];
    keys[i] = keys[j];
    keys[j] = ki;
    const uint32_t ii = ids[i];
    ids[i] = ids[j];
    ids[j] = ii;
  }
  __syncthreads();
  return acc;
}

struct SeqItem {
  uint64_t key;      // location key of a member (the child's pivot)
  uint32_t begin, count;
  uint32_t depth;    // LCA depth (the node's depth)
  uint32_t own;      // the node's own segment length (all of it for a leaf)
  uint32_t nchild;
  uint32_t cfirst;   // BFS id of the first child
};

// One BFS level: every item a block.  Children go to (kid_key, kid_cnt, kid_depth)
// at their first position.
template <int D, bool BOX>
__global__ void __launch_bounds__(kSB) k_seq_level(SeqItem *__restrict__ items, uint64_t *__restrict__ keys,
                                                    uint32_t *__restrict__ ids, uint32_t *__restrict__ ff,
                                                    uint32_t *__restrict__ bt, KeyView kv, uint32_t M, uint32_t cap,
                                                    uint64_t *__restrict__ kid_key, uint32_t *__restrict__ kid_cnt,
                                                    uint8_t *__restrict__ kid_depth) {
  __shared__ uint32_t s_w[kSB / 64];
  __shared__ uint64_t s_r[4 * (kSB / 64)];
  const SeqItem it = items[blockIdx.x];
  const uint32_t b = it.begin, e = it.begin + it.count;
  uint32_t own = it.count, nch = 0;
  if (it.count > cap && it.depth < M) {
    own = 0;
    if (BOX) {
      const SeqAcc a = seq_partition(keys, ids, b, e, SeqPred{true, it.depth, 0, 0, 0}, kv, M, ff, bt, s_w, s_r);
      own = a.count;
    }
    if (own < it.count) {
      const uint32_t shift = (uint32_t)D * (M - (it.depth + 1));
      uint32_t cb = b + own;
      while (cb < e) {
        const uint64_t pivot = keys[cb];
        const uint64_t flag = (kv.code(pivot) >> shift) & ((1ull << D) - 1ull);
        const SeqAcc a = seq_partition(keys, ids, cb, e, SeqPred{false, 0, shift, flag, (1ull << D) - 1ull}, kv, M, ff,
                                       bt, s_w, s_r);
        const uint64_t diff = a.lo ^ a.hi;
        const uint32_t level_diff = (bw64(diff) + D - 1) / D;
        uint32_t lca = M - level_diff;
        if (BOX && a.mind < lca) lca = a.mind;
        if (threadIdx.x == 0) {
          kid_key[cb] = pivot;
          kid_cnt[cb] = a.count;
          kid_depth[cb] = (uint8_t)lca;
        }
        ++nch;
        cb += a.count;
      }
    } else {
      own = it.count;  // only entities of the node's depth: a leaf (:460-465)
    }
  }
  if (threadIdx.x == 0) {
    items[blockIdx.x].own = own;
    items[blockIdx.x].nchild = nch;
  }
}

// next level: the recorded children in position order (clears the records)
__global__ void __launch_bounds__(kB) k_seq_gather_level(const uint32_t *__restrict__ scan, uint64_t n,
                                                          uint64_t *__restrict__ kid_key, uint32_t *__restrict__ kid_cnt,
                                                          const uint8_t *__restrict__ kid_depth,
                                                          SeqItem *__restrict__ next) {
  const uint64_t p = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (p >= n || !kid_cnt[p]) return;
  SeqItem x;
  x.key = kid_key[p];
  x.begin = (uint32_t)p;
  x.count = kid_cnt[p];
  x.depth = kid_depth[p];
  x.own = 0;
  x.nchild = 0;
  x.cfirst = 0;
  next[scan[p]] = x;
  kid_cnt[p] = 0;
}

__global__ void k_seq_child_flags(const SeqItem *__restrict__ items, uint32_t n, uint32_t *__restrict__ nch) {
  const uint32_t i = blockIdx.x * kB + threadIdx.x;
  if (i < n) nch[i] = items[i].nchild;
}

// node records from the BFS items (node x = item x); cfirst: BFS id of the first child
template <int D, bool BOX>
__global__ void __launch_bounds__(kB) k_seq_finalize(const SeqItem *__restrict__ items, const uint32_t *__restrict__ cfirst,
                                                      uint64_t nn, const SetParams *__restrict__ sets, KeyView kv,
                                                      NodeInfo *__restrict__ info, uint32_t *__restrict__ rend,
                                                      uint8_t *__restrict__ ndepth, uint64_t *__restrict__ nodekey,
                                                      double *__restrict__ nmin, double *__restrict__ nsize) {
  const uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nn) return;
  const SeqItem it = items[x];
  const SetParams &sp = sets[0];
  const uint32_t M = sp.max_depth, depth = it.depth;
  NodeInfo ni;
  ni.ebegin = it.begin;
  ni.ecount = it.own;
  ni.cbegin = it.nchild ? cfirst[x] : 0u;
  ni.ccount = it.nchild;
  info[x] = ni;
  rend[x] = it.begin + it.count;
  ndepth[x] = (uint8_t)depth;
  const uint32_t sh = (M - depth) * D;
  const uint64_t nkey = (1ull << (depth * D)) | ((sh >= 64) ? 0ull : (kv.code(it.key) >> sh));
  nodekey[x] = nkey;
#pragma unroll
  for (int d = 0; d < D; ++d) nsize[x * D + d] = sp.depth_size[(BOX ? 0 : depth) * D + d];
  if (!BOX) {  // InitNodeGeometry (MinPoint), as k_finalize
    uint32_t g[D];
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = 0;
    uint32_t shift = 0;
    for (uint32_t lvl = M - depth; lvl < M; ++lvl)
#pragma unroll
      for (int d = 0; d < D; ++d, ++shift) g[d] |= (uint32_t)((nkey >> shift) & 1ull) << lvl;
    const double half_grid = (double)(1ull << (M - depth)) * 0.5;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double center = ((double)g[d] + half_grid) * sp.deraster[d] + sp.space_min[d];
      nmin[x * D + d] = center - sp.depth_size[(depth + 1) * D + d];
      nsize[x * D + d] = sp.depth_size[depth * D + d];
    }
  }
}

template <int D, bool BOX>
static void build_seq_impl(Tree *t, const double *geom_in, int mem, uint64_t n, const double *box, uint32_t depth,
                           uint32_t cap) {
  Ctx *ctx = t->ctx;
  constexpr int S = BOX ? 2 * D : D;
  t->n = n;
  if (n >= (1ull << 30)) throw Error{OTG_ERR_UNSUPPORTED, "more than 2^30 entities"};
  Buf<double> dgeom_own;
  const double *dgeom = geom_in;
  if (mem == OTG_MEM_HOST) {
    dgeom_own = Buf<double>(ctx, n * S);
    if (n) OTG_HIP_CHECK(hipMemcpyAsync(dgeom_own.p, geom_in, n * S * 8, hipMemcpyHostToDevice, ctx->stream));
    dgeom = dgeom_own.p;
  }
  const uint64_t offs[2] = {0, n};
  prepare_sets<D, BOX>(t, dgeom, offs, 1, box, depth ? &depth : nullptr, &cap);
  Buf<uint64_t> d_off(ctx, 2);
  OTG_HIP_CHECK(hipMemcpyAsync(d_off.p, offs, 16, hipMemcpyHostToDevice, ctx->stream));
  upload_sets(t);
  if (n == 0) {
    t->node_count = 0;
    return;
  }
  const SetParams &sp = t->sets[0];
  const uint32_t M = sp.max_depth;
  const KeyView kv{t->layout.dbits, t->layout.cbits};
  // locations in input order; the entity order the partitions evolve
  t->d_keys = (uint64_t *)ctx->alloc(n * 8);
  Buf<uint32_t> order(ctx, n), err(ctx, 1);
  OTG_HIP_CHECK(hipMemsetAsync(err.p, 0, 4, ctx->stream));
  launch(ctx, "encode", k_encode<D, BOX>, dim3(div_up(n, kB * kPointItems<BOX ? 2 * D : D>)), dim3(kB), 0, dgeom, n, (const uint64_t *)d_off.p, 1u,
         (const SetParams *)t->d_sets, kv, t->d_keys, order.p, err.p, 0u);
  Buf<uint32_t> ff(ctx, n), bt(ctx, n), kid_cnt(ctx, n), scan(ctx, n + 1), total(ctx, 1);
  Buf<uint64_t> kid_key(ctx, n);
  Buf<uint8_t> kid_depth(ctx, n);
  OTG_HIP_CHECK(hipMemsetAsync(kid_cnt.p, 0, n * 4, ctx->stream));
  // BFS levels, concatenated: node x = item x
  std::vector<SeqItem> root(1);
  root[0] = SeqItem{0, 0, (uint32_t)n, 0, 0, 0, 0};
  std::vector<uint64_t> level_off{0, 1};
  uint64_t cap_items = std::max<uint64_t>(1024, n / 4);
  Buf<SeqItem> items(ctx, cap_items);
  OTG_HIP_CHECK(hipMemcpyAsync(items.p, root.data(), sizeof(SeqItem), hipMemcpyHostToDevice, ctx->stream));
  uint32_t herr = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  if (herr & kErrOutOfSpace) throw Error{OTG_ERR_OUT_OF_SPACE, "point outside the given box"};
  if (herr & kErrUnsupported) throw Error{OTG_ERR_UNSUPPORTED, "box level or grid id beyond the tree depth"};
  for (;;) {
    const uint64_t l0 = level_off[level_off.size() - 2], l1 = level_off.back();
    launch(ctx, "seq_level", k_seq_level<D, BOX>, dim3((uint32_t)(l1 - l0)), dim3(kSB), 0, items.p + l0, t->d_keys,
           order.p, ff.p, bt.p, kv, M, sp.cap, kid_key.p, kid_cnt.p, kid_depth.p);
    // children recorded at their first position, in position order
    Buf<uint32_t> flags(ctx, n);
    launch(ctx, "seq_flags", k_nonzero_u32, dim3(div_up(n, kB)), dim3(kB), 0, (const uint32_t *)kid_cnt.p, n, flags.p);
    exclusive_scan_u32(ctx, flags.p, scan.p, n, total.p);
    uint32_t next = 0;
    OTG_HIP_CHECK(hipMemcpyAsync(&next, total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (next == 0) break;
    if (l1 + next > cap_items) {
      Buf<SeqItem> bigger(ctx, std::max<uint64_t>(2 * cap_items, l1 + next));
      OTG_HIP_CHECK(hipMemcpyAsync(bigger.p, items.p, l1 * sizeof(SeqItem), hipMemcpyDeviceToDevice, ctx->stream));
      std::swap(items, bigger);
      cap_items = std::max<uint64_t>(2 * cap_items, l1 + next);
    }
    launch(ctx, "seq_gather_level", k_seq_gather_level, dim3(div_up(n, kB)), dim3(kB), 0, (const uint32_t *)scan.p, n,
           kid_key.p, kid_cnt.p, (const uint8_t *)kid_depth.p, items.p + l1);
    level_off.push_back(l1 + next);
  }
  const uint64_t nn = level_off.back();
  // first child of each node: the children of level L are level L + 1 in parent order
  Buf<uint32_t> nch(ctx, nn), cfirst(ctx, nn);
  launch(ctx, "seq_child_counts", k_seq_child_flags, dim3(div_up(nn, kB)), dim3(kB), 0, (const SeqItem *)items.p,
         (uint32_t)nn, nch.p);
  exclusive_scan_u32(ctx, nch.p, cfirst.p, nn, total.p);
  launch(ctx, "seq_child_base", k_add_u32, dim3(div_up(nn, kB)), dim3(kB), 0, cfirst.p, nn, 1u);  // + the root
  t->node_count = nn;
  t->sets[0].node_off = 0;
  t->sets[0].node_cnt = nn;
  t->d_info = (NodeInfo *)ctx->alloc(nn * sizeof(NodeInfo));
  t->d_rend = (uint32_t *)ctx->alloc(nn * 4);
  t->d_depth = (uint8_t *)ctx->alloc(nn);
  t->d_nkey = (uint64_t *)ctx->alloc(nn * 8);
  t->d_nmin = (double *)ctx->alloc(nn * D * 8);
  t->d_nsize = (double *)ctx->alloc(nn * D * 8);
  upload_sets(t);
  launch(ctx, "seq_finalize", k_seq_finalize<D, BOX>, dim3(div_up(nn, kB)), dim3(kB), 0, (const SeqItem *)items.p,
         (const uint32_t *)cfirst.p, nn, (const SetParams *)t->d_sets, kv, t->d_info, t->d_rend, t->d_depth, t->d_nkey,
         t->d_nmin, t->d_nsize);
  // entities in the order the partitions left them
  t->d_geom = (double *)ctx->alloc(n * S * 8);
  t->d_ids = (uint32_t *)ctx->alloc(n * 4);
  launch(ctx, "gather", k_gather<D, BOX>, dim3(div_up(n, kB)), dim3(kB), 0, dgeom, (const uint32_t *)order.p, n,
         (const uint64_t *)d_off.p, 1u, t->d_geom, t->d_ids);
  if (BOX) build_mbr<D>(t);
  SeqItem r0;
  OTG_HIP_CHECK(hipMemcpyAsync(&r0, items.p, sizeof(SeqItem), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  t->sets[0].root_splits = r0.nchild ? 1u : 0u;
  upload_sets(t);
  t->device_bytes = n * (S * 8 + 4 + 8) + nn * (sizeof(NodeInfo) + 4 + 1 + 8 + 2 * D * 8);
  // points: the finer kNN / range index when max-depth leaves overflow, as a build makes it
  if (!BOX) {
    const uint32_t fmax = finer_depth(D, 0);
    if (M < fmax) {
      Buf<uint32_t> flag(ctx, 1);
      OTG_HIP_CHECK(hipMemsetAsync(flag.p, 0, 4, ctx->stream));
      launch(ctx, "oversized", k_oversized, dim3(div_up(nn, kB)), dim3(kB), 0, (const NodeInfo *)t->d_info, nn,
             (const SetParams *)t->d_sets, 1u, flag.p);
      uint32_t hflag = 0;
      OTG_HIP_CHECK(hipMemcpyAsync(&hflag, flag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
      if (hflag) {
        const uint32_t fdepth = finer_set_depth(t->sets[0], D, fmax);
        uint32_t icap = ctx->tuning.index_capacity ? ctx->tuning.index_capacity : cap;
        double sbox[2 * kMaxDim];
        for (int d = 0; d < D; ++d) {
          sbox[d] = t->sets[0].space_min[d];
          sbox[D + d] = t->sets[0].space_max[d];
        }
        Tree *x = new Tree();
        x->ctx = ctx;
        x->kind = t->kind;
        x->dim = t->dim;
        try {
          build_impl<D, false>(x, dgeom, OTG_MEM_DEVICE, offs, 1, sbox, &fdepth, &icap, false);
        } catch (...) {
          delete x;
          throw;
        }
        t->knn_tree = x;
        t->device_bytes += x->device_bytes;
      }
    }
  }
}

void build_seq(Tree *t, const double *geom, int mem, uint64_t n, const double *box, uint32_t depth, uint32_t cap) {
  const bool box_kind = t->kind == OTG_KIND_BOX;
  switch (t->dim) {
#define OTG_CASE(DD)                                                          \
  case DD:                                                                    \
    if (box_kind) build_seq_impl<DD, true>(t, geom, mem, n, box, depth, cap); \
    else build_seq_impl<DD, false>(t, geom, mem, n, box, depth, cap);         \
    break;
    OTG_CASE(2)
    OTG_CASE(3)
    OTG_CASE(4)
    OTG_CASE(5)
    OTG_CASE(6)
#undef OTG_CASE
    default:
      throw Error{OTG_ERR_INVALID_ARG, "dimension must be 2..6"};
  }
  publish_views(t);
}

// ---------------------------------------------------------------------------
// loading a static-core archive: serialize(BinaryInputArchive&,
// StaticLinearOrthoTreeCore&) (serialization.h:359-386, binary_archive.h) reads the
// members otg_tree_serialize writes, in the same order
// ---------------------------------------------------------------------------
struct ArchiveIn {
  const uint8_t *p;
  uint64_t n, pos;
  bool large;
  template <typename T>
  T get() {
    if (n - pos < sizeof(T)) throw Error{OTG_ERR_INVALID_ARG, "the archive is truncated"};
    T v;
    std::memcpy(&v, p + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  uint64_t count() { return large ? get<uint64_t>() : (uint64_t)get<uint32_t>(); }
  void vec(double *out, int D) {  // std::array<double, D>: its size, then the values
    if (count() != (uint64_t)D) throw Error{OTG_ERR_INVALID_ARG, "archive array length is not the dimension"};
    for (int d = 0; d < D; ++d) out[d] = get<double>();
  }
  // count of a vector whose elements take at least `min_bytes` each
  uint64_t count_of(uint64_t min_bytes) {
    const uint64_t c = count();
    if (min_bytes && c > (n - pos) / min_bytes) throw Error{OTG_ERR_INVALID_ARG, "archive count exceeds its bytes"};
    return c;
  }
};

struct LoadedTable {
  uint64_t cap = 0;
  uint32_t M = 0;
  double box_min[kMaxDim] = {}, box_size[kMaxDim] = {};
  std::vector<double> depth_size;  // points: m_nodeSize, (M + 2) x D
  std::vector<uint64_t> begins, eb, ec;
  std::vector<uint8_t> depth;
  std::vector<uint32_t> ids;
  std::vector<double> nmin, nsize;
  std::vector<double> entities;  // OrthoTreeManaged archives: the entity vector after the core
};

template <typename Index, typename Seg>
static void read_node_storage(ArchiveIn &ar, LoadedTable &L) {
  const uint64_t nb = ar.count_of(sizeof(Index));
  L.begins.resize(nb);
  for (uint64_t i = 0; i < nb; ++i) L.begins[i] = ar.get<Index>();
  const uint64_t nn = ar.count_of(2 * sizeof(Seg));
  L.eb.resize(nn);
  L.ec.resize(nn);
  for (uint64_t i = 0; i < nn; ++i) {
    L.eb[i] = ar.get<Seg>();
    L.ec[i] = ar.get<Seg>();
  }
}

static LoadedTable read_static_archive(ArchiveIn &ar, int D, bool box, bool want_entities) {
  LoadedTable L;
  if (ar.get<uint32_t>() != 0 || ar.get<uint32_t>() != 0)
    throw Error{OTG_ERR_INVALID_ARG, "archive version is not StaticLinearOrthoTreeCore 0 / OrthoTreeCoreBase 0"};
  L.cap = ar.get<uint64_t>();
  L.M = ar.get<uint32_t>();
  if (L.cap == 0 || L.M < 1 || L.M > (uint32_t)kMaxDepth || L.M > max_theoretical_depth(D))
    throw Error{OTG_ERR_INVALID_ARG, "archive capacity or depth out of range"};
  if (box) {
    ar.vec(L.box_min, D);
    ar.vec(L.box_size, D);
  } else {
    if (ar.count() != (uint64_t)L.M + 2) throw Error{OTG_ERR_INVALID_ARG, "archive node sizes are not depth + 2"};
    L.depth_size.resize((size_t)(L.M + 2) * D);
    for (uint32_t k = 0; k < L.M + 2; ++k) ar.vec(&L.depth_size[(size_t)k * D], D);
  }
  const uint64_t kind = ar.count();
  if (kind == 0) read_node_storage<uint8_t, uint8_t>(ar, L);
  else if (kind == 1) read_node_storage<uint16_t, uint16_t>(ar, L);
  else if (kind == 2 && ar.large) read_node_storage<uint32_t, uint64_t>(ar, L);
  else if (kind == 2) read_node_storage<uint32_t, uint32_t>(ar, L);
  else throw Error{OTG_ERR_INVALID_ARG, "archive node storage index is not 0, 1 or 2"};
  const uint64_t nn = L.eb.size();
  if (ar.count_of(1) != nn) throw Error{OTG_ERR_INVALID_ARG, "archive depth count differs from the node count"};
  L.depth.resize(nn);
  for (uint64_t i = 0; i < nn; ++i) L.depth[i] = ar.get<uint8_t>();
  const uint64_t ne = ar.count_of(4);
  L.ids.resize(ne);
  for (uint64_t i = 0; i < ne; ++i) L.ids[i] = ar.get<uint32_t>();
  if (ar.count_of(8) != nn) throw Error{OTG_ERR_INVALID_ARG, "archive geometry count differs from the node count"};
  L.nmin.resize(nn * D);
  if (box) L.nsize.resize(nn * D);
  for (uint64_t i = 0; i < nn; ++i) {
    ar.vec(&L.nmin[i * D], D);
    if (box) ar.vec(&L.nsize[i * D], D);
  }
  if (want_entities) {
    const int S = box ? 2 * D : D;
    if (ar.count_of(8) != ne) throw Error{OTG_ERR_INVALID_ARG, "archive entity vector differs from the entity count"};
    L.entities.resize(ne * S);
    for (uint64_t i = 0; i < ne; ++i) {
      ar.vec(&L.entities[i * S], D);
      if (box) ar.vec(&L.entities[i * S + D], D);
    }
  }
  if (ar.pos != ar.n) throw Error{OTG_ERR_INVALID_ARG, "archive has bytes after the tree"};
  return L;
}

// Node keys of a loaded box tree: the code of the node's first subtree entity at
// the node's depth (the node's cell), with the depth's leading bit (si_morton.h:216-226);
// also the sorted keys of the storage order.
__global__ void __launch_bounds__(kB) k_loaded_keys(const uint64_t *__restrict__ keys_in, const uint32_t *__restrict__ ids,
                                                     uint64_t n, uint64_t *__restrict__ keys_sorted) {
  const uint64_t j = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (j < n) keys_sorted[j] = keys_in[ids[j]];
}

template <int D>
__global__ void __launch_bounds__(kB) k_loaded_node_keys(const NodeInfo *__restrict__ info, const uint8_t *__restrict__ depth,
                                                          uint64_t nn, const uint64_t *__restrict__ keys_sorted,
                                                          uint64_t ne, KeyView kv, uint32_t M,
                                                          uint64_t *__restrict__ nodekey) {
  const uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nn) return;
  const uint32_t d = depth[x];
  const uint64_t code = info[x].ebegin < ne ? kv.code(keys_sorted[info[x].ebegin]) : 0ull;
  const uint32_t sh = (M - d) * D;
  nodekey[x] = (1ull << (d * D)) | ((sh >= 64) ? 0ull : (code >> sh));
}

// Entities of each node's own segment lie in the node's box (MinPoint box for
// points, MBR for boxes), to a relative 1e-9: the archive was written for these entities.
template <int D, bool BOX>
__global__ void __launch_bounds__(kB) k_loaded_check(const NodeInfo *__restrict__ info, uint64_t nn,
                                                      const double *__restrict__ geom, uint64_t n,
                                                      const double *__restrict__ nmin, const double *__restrict__ nsize,
                                                      uint32_t *__restrict__ bad) {
  const uint64_t x = (uint64_t)blockIdx.x * kB + threadIdx.x;
  if (x >= nn) return;
  const NodeInfo ni = info[x];
  bool ok = true;
  for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount && ok; ++j)
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double lo = nmin[x * D + d], hi = lo + nsize[x * D + d];
      const double eps = 1e-9 * (fabs(lo) + fabs(hi) + 1.0);
      const double a = geom[(uint64_t)d * n + j], b = BOX ? geom[(uint64_t)(D + d) * n + j] : a;
      ok &= (a >= lo - eps) && (b <= hi + eps);
    }
  if (!ok) atomicOr(bad, 1u);
}

template <int D, bool BOX>
static void load_impl(Tree *t, const uint8_t *buf, uint64_t size, bool large, uint64_t n, const double *geom_in,
                      int mem) {
  Ctx *ctx = t->ctx;
  constexpr int S = BOX ? 2 * D : D;
  ArchiveIn ar{buf, size, 0, large};
  LoadedTable L = read_static_archive(ar, D, BOX, geom_in == nullptr);
  const uint64_t nn = L.eb.size(), ne = L.ids.size();
  if (geom_in && n != ne) throw Error{OTG_ERR_INVALID_ARG, "the entity count differs from the archive's"};
  if (ne >= (1ull << 30)) throw Error{OTG_ERR_UNSUPPORTED, "more than 2^30 entities"};
  if ((nn == 0) != (ne == 0)) throw Error{OTG_ERR_INVALID_ARG, "archive has entities but no nodes, or nodes but no entities"};

  // structure: ids a permutation; children after their parent, deeper, within the
  // table; the static linear layout (a node's own entities, then each child's
  // subtree, contiguous), which also bounds every walk the queries make
  std::vector<uint8_t> seen(ne, 0);
  for (uint32_t id : L.ids) {
    if (id >= ne || seen[id]) throw Error{OTG_ERR_INVALID_ARG, "archive entity ids are not a permutation"};
    seen[id] = 1;
  }
  std::vector<NodeInfo> info(nn);
  std::vector<uint32_t> rend(nn);
  if (nn) {
    if (L.begins.empty() || L.begins[0] != 1 || L.begins.size() > nn + 1)
      throw Error{OTG_ERR_INVALID_ARG, "archive child segments are malformed"};
    for (uint64_t i = 0; i < nn; ++i) {
      NodeInfo &ni = info[i];
      if (L.depth[i] > L.M || L.eb[i] > ne || L.ec[i] > ne - L.eb[i])
        throw Error{OTG_ERR_INVALID_ARG, "archive node depth or entity segment out of range"};
      ni.ebegin = (uint32_t)L.eb[i];
      ni.ecount = (uint32_t)L.ec[i];
      ni.cbegin = 0;
      ni.ccount = 0;
      if (i + 1 < L.begins.size()) {
        if (L.begins[i + 1] < L.begins[i] || L.begins[i + 1] > nn)
          throw Error{OTG_ERR_INVALID_ARG, "archive child segments are malformed"};
        if (L.begins[i + 1] > L.begins[i]) {
          if (L.begins[i] <= i) throw Error{OTG_ERR_INVALID_ARG, "archive children precede their parent"};
          ni.cbegin = (uint32_t)L.begins[i];
          ni.ccount = (uint32_t)(L.begins[i + 1] - L.begins[i]);
        }
      }
    }
    if (L.depth[0] != 0) throw Error{OTG_ERR_INVALID_ARG, "archive root is not at depth 0"};
    std::vector<uint8_t> has_parent(nn, 0);
    for (uint64_t i = nn; i-- > 0;) {
      const NodeInfo &ni = info[i];
      uint64_t pos = (uint64_t)ni.ebegin + ni.ecount;
      for (uint32_t c = ni.cbegin; c < ni.cbegin + ni.ccount; ++c) {
        if (L.depth[c] <= L.depth[i] || has_parent[c] || info[c].ebegin != pos)
          throw Error{OTG_ERR_INVALID_ARG, "archive tree is not a static linear layout"};
        has_parent[c] = 1;
        pos = rend[c];
      }
      rend[i] = (uint32_t)pos;
    }
    for (uint64_t i = 1; i < nn; ++i)
      if (!has_parent[i]) throw Error{OTG_ERR_INVALID_ARG, "archive node without a parent"};
    if (info[0].ebegin != 0 || rend[0] != ne) throw Error{OTG_ERR_INVALID_ARG, "archive root does not hold every entity"};
  }

  // set parameters: the tree box of a box tree is stored; a point tree stores its
  // node sizes, and its root MinPoint stands in for the box minimum (used only to
  // order query batches)
  t->n = ne;
  t->node_count = nn;
  t->sets.assign(1, SetParams{});
  SetParams &sp = t->sets[0];
  sp.ent_off = 0;
  sp.ent_cnt = ne;
  sp.node_off = 0;
  sp.node_cnt = nn;
  sp.max_depth = L.M;
  sp.cap = (uint32_t)std::min<uint64_t>(L.cap, 0xFFFFFFFFull);
  sp.min_lcp = 255;
  sp.root_splits = (nn && info[0].ccount) ? 1u : 0u;
  {
    volatile double res = (double)(1ull << L.M);
    for (int d = 0; d < D; ++d) {
      const double mn = BOX ? L.box_min[d] : (nn ? L.nmin[d] : 0.0);
      volatile double sz = BOX ? L.box_size[d] : L.depth_size[d];
      sp.space_min[d] = mn;
      sp.space_max[d] = mn + sz;
      sp.size[d] = sz;
      sp.deraster[d] = sz / res;
      sp.raster[d] = (sz == 0) ? 1.0 : (res / sz);
    }
    if (BOX) {
      double loose = 2.0;
      for (uint32_t k = 0; k < L.M + 2; ++k, loose *= 0.5)
        for (int d = 0; d < D; ++d) sp.depth_size[k * D + d] = sp.size[d] * loose;
    } else {
      std::copy(L.depth_size.begin(), L.depth_size.end(), sp.depth_size);
    }
  }
  set_layout<D, BOX>(t);
  upload_sets(t);
  if (ne == 0) return;

  // node arrays
  std::vector<uint64_t> nkey(nn, 0);
  std::vector<double> nsize(nn * D);
  for (uint64_t i = 0; i < nn; ++i)
    for (int d = 0; d < D; ++d) nsize[i * D + d] = BOX ? L.nsize[i * D + d] : L.depth_size[(size_t)L.depth[i] * D + d];
  if (!BOX)  // the node's cell at its depth, from its MinPoint relative to the root's
    for (uint64_t i = 0; i < nn; ++i) {
      const uint32_t dep = L.depth[i];
      uint32_t g[D];
      for (int d = 0; d < D; ++d) {
        const double cell = L.depth_size[(size_t)dep * D + d];
        const double x = cell > 0 ? std::nearbyint((L.nmin[i * D + d] - L.nmin[d]) / cell) : 0.0;
        g[d] = (uint32_t)std::min<double>(std::max(x, 0.0), (double)((1ull << dep) - 1));
      }
      uint64_t code = 0;
      for (uint32_t b = 0; b < dep; ++b)
        for (int d = 0; d < D; ++d) code |= (uint64_t)((g[d] >> b) & 1u) << (b * D + d);
      nkey[i] = (1ull << (dep * D)) | code;
    }
  t->d_info = (NodeInfo *)ctx->alloc(nn * sizeof(NodeInfo));
  t->d_rend = (uint32_t *)ctx->alloc(nn * 4);
  t->d_depth = (uint8_t *)ctx->alloc(nn);
  t->d_nkey = (uint64_t *)ctx->alloc(nn * 8);
  t->d_nmin = (double *)ctx->alloc(nn * D * 8);
  t->d_nsize = (double *)ctx->alloc(nn * D * 8);
  t->d_ids = (uint32_t *)ctx->alloc(ne * 4);
  t->d_geom = (double *)ctx->alloc(ne * S * 8);
  hipStream_t st = ctx->stream;
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_info, info.data(), nn * sizeof(NodeInfo), hipMemcpyHostToDevice, st));
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_rend, rend.data(), nn * 4, hipMemcpyHostToDevice, st));
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_depth, L.depth.data(), nn, hipMemcpyHostToDevice, st));
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_nkey, nkey.data(), nn * 8, hipMemcpyHostToDevice, st));
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_nmin, L.nmin.data(), nn * D * 8, hipMemcpyHostToDevice, st));
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_nsize, nsize.data(), nn * D * 8, hipMemcpyHostToDevice, st));
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_ids, L.ids.data(), ne * 4, hipMemcpyHostToDevice, st));

  // entities (input order) on the device, then in storage order
  Buf<double> dgeom_own;
  const double *dgeom = geom_in;
  if (!geom_in || mem == OTG_MEM_HOST) {
    dgeom_own = Buf<double>(ctx, ne * S);
    OTG_HIP_CHECK(hipMemcpyAsync(dgeom_own.p, geom_in ? geom_in : L.entities.data(), ne * S * 8,
                                 hipMemcpyHostToDevice, st));
    dgeom = dgeom_own.p;
  }
  const KeyView kv{t->layout.dbits, t->layout.cbits};
  Buf<uint64_t> d_off(ctx, 2);
  const uint64_t offs[2] = {0, ne};
  OTG_HIP_CHECK(hipMemcpyAsync(d_off.p, offs, 16, hipMemcpyHostToDevice, st));
  Buf<uint32_t> ids_copy(ctx, ne);
  OTG_HIP_CHECK(hipMemcpyAsync(ids_copy.p, t->d_ids, ne * 4, hipMemcpyDeviceToDevice, st));
  launch(ctx, "gather", k_gather<D, BOX>, dim3(div_up(ne, kB)), dim3(kB), 0, dgeom, (const uint32_t *)ids_copy.p, ne,
         (const uint64_t *)d_off.p, 1u, t->d_geom, t->d_ids);
  Buf<uint32_t> err(ctx, 2);
  OTG_HIP_CHECK(hipMemsetAsync(err.p, 0, 8, st));
  launch(ctx, "loaded_check", k_loaded_check<D, BOX>, dim3(div_up(nn, kB)), dim3(kB), 0, (const NodeInfo *)t->d_info,
         nn, (const double *)t->d_geom, ne, (const double *)t->d_nmin, (const double *)t->d_nsize, err.p);
  if (BOX) {  // location keys against the stored tree box, then node keys
    Buf<uint64_t> keys(ctx, ne);
    Buf<uint32_t> vals(ctx, ne);
    launch(ctx, "encode", k_encode<D, true>, dim3(div_up(ne, kB * kPointItems<2 * D>)), dim3(kB), 0, dgeom, ne, (const uint64_t *)d_off.p,
           1u, (const SetParams *)t->d_sets, kv, keys.p, vals.p, err.p + 1, 0u);
    t->d_keys = (uint64_t *)ctx->alloc(ne * 8);
    launch(ctx, "loaded_keys", k_loaded_keys, dim3(div_up(ne, kB)), dim3(kB), 0, (const uint64_t *)keys.p,
           (const uint32_t *)t->d_ids, ne, t->d_keys);
    launch(ctx, "loaded_node_keys", k_loaded_node_keys<D>, dim3(div_up(nn, kB)), dim3(kB), 0,
           (const NodeInfo *)t->d_info, (const uint8_t *)t->d_depth, nn, (const uint64_t *)t->d_keys, ne, kv, L.M,
           t->d_nkey);
  }
  uint32_t herr[2] = {0, 0};
  OTG_HIP_CHECK(hipMemcpyAsync(herr, err.p, 8, hipMemcpyDeviceToHost, st));
  ctx->sync();
  if (herr[0]) throw Error{OTG_ERR_INVALID_ARG, "an entity lies outside its node in the archive: not these entities"};
  if (herr[1]) throw Error{OTG_ERR_INVALID_ARG, "an entity lies outside the archive's tree box"};
  t->device_bytes = ne * (S * 8 + 4 + (BOX ? 8 : 0)) + nn * (sizeof(NodeInfo) + 4 + 1 + 8 + 2 * D * 8);

  // points: the finer kNN / range index when max-depth leaves overflow, as a build makes it
  if (!BOX) {
    bool overfull = false;
    for (uint64_t i = 0; i < nn && !overfull; ++i) overfull = !info[i].ccount && info[i].ecount > sp.cap;
    const uint32_t fmax = finer_depth(D, 0);
    if (overfull && sp.max_depth < fmax) {
      const uint32_t fdepth = finer_set_depth(sp, D, fmax);
      uint32_t icap = ctx->tuning.index_capacity ? ctx->tuning.index_capacity : sp.cap;
      Tree *x = new Tree();
      x->ctx = ctx;
      x->kind = t->kind;
      x->dim = t->dim;
      try {
        build_impl<D, false>(x, dgeom, OTG_MEM_DEVICE, offs, 1, nullptr, &fdepth, &icap, false);
      } catch (...) {
        delete x;
        throw;
      }
      t->knn_tree = x;
      t->device_bytes += x->device_bytes;
    }
  }
}

void load_static_archive(Tree *t, const uint8_t *buf, uint64_t size, bool large, uint64_t n, const double *geom,
                         int mem) {
  const bool box = t->kind == OTG_KIND_BOX;
  switch (t->dim) {
#define OTG_CASE(DD)                                               \
  case DD:                                                         \
    if (box) load_impl<DD, true>(t, buf, size, large, n, geom, mem);  \
    else load_impl<DD, false>(t, buf, size, large, n, geom, mem);     \
    break;
    OTG_CASE(2)
    OTG_CASE(3)
    OTG_CASE(4)
    OTG_CASE(5)
    OTG_CASE(6)
#undef OTG_CASE
    default:
      throw Error{OTG_ERR_INVALID_ARG, "dimension must be 2..6"};
  }
  publish_views(t);
}

// ---------------------------------------------------------------------------
// build plan: a point-forest build captured once as a HIP graph and replayed on new
// contents of the same device geometry (same sets, boxes, depths and capacities)
// ---------------------------------------------------------------------------
// Words the captured build leaves on the device (and copies to pinned memory last):
// the counts of each node table, a capacity overflow flag and the encode error bits.
constexpr int kPlanRefCounts = 0, kPlanFineCounts = 4, kPlanOverflow = 8, kPlanErr = 9, kPlanWords = 16;

struct BuildPlan {
  Ctx *ctx = nullptr;
  std::unique_ptr<::otg_tree> t;  // the plan's tree (an ABI handle); its arrays live in `mem`
  const double *geom = nullptr;
  uint64_t n = 0;
  int dim = 0;
  bool fine = false;        // a finer kNN/range index sorted together with the tree
  char *mem = nullptr;      // every device buffer of the captured build
  TableSizes ref, fine_sz;  // capacities and device counts of the two node tables
  uint64_t cap_ref_nodes = 0;
  uint32_t *d_words = nullptr;
  uint32_t *h_words = nullptr;  // pinned
  SetParams *h_init = nullptr;  // pinned: sets as each replay starts them (tree, then finer index)
  SetParams *h_sets = nullptr;  // pinned: sets as the last replay left them
  uint64_t *h_off = nullptr;    // pinned set offsets
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BuildPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    t.reset();
    for (void *h : {(void *)h_words, (void *)h_init, (void *)h_sets, (void *)h_off})
      if (h) (void)hipHostFree(h);
    if (ctx && mem) ctx->free(mem);
  }
};

// The captured work: every buffer from `ar`, every count kept on the device.
template <int D>
static void build_plan_body(BuildPlan &p, Arena &ar) {
  Ctx *ctx = p.ctx;
  Tree *t = p.t.get();
  Tree *aux = t->knn_tree;
  const uint32_t nsets = (uint32_t)t->sets.size();
  const uint64_t n = p.n;
  fill_u32(ctx, p.d_words, kPlanWords, 0u);
  OTG_HIP_CHECK(hipMemcpyAsync(t->d_sets, p.h_init, sizeof(SetParams) * nsets, hipMemcpyHostToDevice, ctx->stream));
  if (aux)
    OTG_HIP_CHECK(hipMemcpyAsync(aux->d_sets, p.h_init + nsets, sizeof(SetParams) * nsets, hipMemcpyHostToDevice,
                                 ctx->stream));
  Buf<uint64_t> d_off(&ar, nsets + 1);
  OTG_HIP_CHECK(hipMemcpyAsync(d_off.p, p.h_off, (nsets + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
  Buf<uint32_t> vals(&ar, n);
  uint32_t *err = p.d_words + kPlanErr;
  if (aux) {
    encode_sort<D, false>(aux, p.geom, d_off.p, vals.p, err, &ar);
    t->d_keys = (uint64_t *)ar.take(n * 8);
    launch(ctx, "coarse_keys", k_coarse_keys<D>, dim3(div_up(n, kB)), dim3(kB), 0, (const uint64_t *)aux->d_keys, n,
           KeyView{aux->layout.dbits, aux->layout.cbits}, (const SetParams *)aux->d_sets,
           KeyView{t->layout.dbits, t->layout.cbits}, (const SetParams *)t->d_sets, t->d_keys);
    node_table<D, false>(aux, d_off.p, nullptr, &ar, &p.fine_sz);
    reference_from_finer<D>(t, aux, &ar, &p.fine_sz, &p.ref);
    gather_entities<D, false>(t, p.geom, vals.p, d_off.p, &ar);
    aux->d_geom = t->d_geom;
    aux->d_ids = t->d_ids;
  } else {
    encode_sort<D, false>(t, p.geom, d_off.p, vals.p, err, &ar);
    node_table<D, false>(t, d_off.p, nullptr, &ar, &p.ref);
    gather_entities<D, false>(t, p.geom, vals.p, d_off.p, &ar);
  }
  OTG_HIP_CHECK(hipMemcpyAsync(p.h_sets, t->d_sets, sizeof(SetParams) * nsets, hipMemcpyDeviceToHost, ctx->stream));
  if (aux)
    OTG_HIP_CHECK(hipMemcpyAsync(p.h_sets + nsets, aux->d_sets, sizeof(SetParams) * nsets, hipMemcpyDeviceToHost,
                                 ctx->stream));
  OTG_HIP_CHECK(hipMemcpyAsync(p.h_words, p.d_words, kPlanWords * 4, hipMemcpyDeviceToHost, ctx->stream));
}

// Device bytes the captured build takes (temporaries are not reused inside an arena).
static size_t build_plan_bytes(const BuildPlan &p, uint32_t nsets) {
  const uint64_t n = p.n, D = (uint64_t)p.dim;
  const uint64_t tiles = n / 4096 + 2;
  const auto table = [&](const TableSizes &ts) {  // one node table: per-position passes and node buffers
    return n * 24 + tiles * 64 + (ts.cap_seg + 2) * 12 + ts.cap_nodes * (8 + 4 + 12 + 29 + 16 * D) +
           (ts.cap_nodes / 4096 + 2) * 2048 + (1u << 20);
  };
  size_t b = n * (4 + 8 + 12) + tiles * 2048 + n * (8 * D + 4) + nsets * 16 + (4u << 20);
  b += table(p.ref);
  if (p.fine) b += n * 8 + table(p.fine_sz) + p.fine_sz.cap_nodes * 8 + p.cap_ref_nodes * (29 + 16 * D);
  return b + b / 8;
}

static uint64_t plan_cap(uint64_t count) { return count + count / 8 + 4096; }

template <int D>
static void build_plan_capture(BuildPlan &p, const Tree &eager, const uint64_t *offsets, uint32_t nsets,
                               const double *boxes, const uint32_t *depths, const uint32_t *caps) {
  Ctx *ctx = p.ctx;
  // the plan's tree: the same sets as the plain build, arrays in the arena
  Tree *t = p.t.get();
  prepare_sets<D, false>(t, nullptr, offsets, nsets, boxes, depths, caps);
  t->n = p.n;
  t->plan_arrays = true;
  const Tree *ef = eager.knn_tree;
  p.fine = ef != nullptr;
  p.ref.cap_seg = plan_cap(p.fine ? 0 : eager.seg_count);
  p.ref.cap_nodes = plan_cap(eager.node_count);
  p.cap_ref_nodes = p.ref.cap_nodes;
  if (p.fine) {
    Tree *aux = new_finer_tree(t, finer_depth(D, t->layout.sbits));
    set_layout<D, false>(aux);
    aux->plan_arrays = true;
    aux->shares_entities = true;
    t->knn_tree = aux;
    t->fine_order = true;
    p.fine_sz.cap_seg = plan_cap(ef->seg_count);
    p.fine_sz.cap_nodes = plan_cap(ef->node_count);
    for (uint32_t q = 0; q < nsets; ++q)
      if (aux->sets[q].cap != t->sets[q].cap)
        throw Error{OTG_ERR_UNSUPPORTED, "build plans need the finer index at the tree's capacity"};
  }
  OTG_HIP_CHECK(hipHostMalloc((void **)&p.h_words, kPlanWords * 4, hipHostMallocDefault));
  OTG_HIP_CHECK(hipHostMalloc((void **)&p.h_init, sizeof(SetParams) * 2 * nsets, hipHostMallocDefault));
  OTG_HIP_CHECK(hipHostMalloc((void **)&p.h_sets, sizeof(SetParams) * 2 * nsets, hipHostMallocDefault));
  OTG_HIP_CHECK(hipHostMalloc((void **)&p.h_off, (nsets + 1) * 8, hipHostMallocDefault));
  std::memset(p.h_words, 0, kPlanWords * 4);
  std::memcpy(p.h_init, t->sets.data(), sizeof(SetParams) * nsets);
  if (p.fine) std::memcpy(p.h_init + nsets, t->knn_tree->sets.data(), sizeof(SetParams) * nsets);
  std::memcpy(p.h_off, offsets, (nsets + 1) * 8);
  for (size_t bytes = build_plan_bytes(p, nsets);; bytes *= 2) {
    p.mem = (char *)ctx->alloc(bytes);
    Arena ar{p.mem, bytes, 0};
    // the words first: no tree array starts at the arena's (allocator-owned) base
    p.d_words = (uint32_t *)ar.take(kPlanWords * 4);
    p.ref.counts = p.d_words + kPlanRefCounts;
    p.ref.overflow = p.d_words + kPlanOverflow;
    p.fine_sz.counts = p.d_words + kPlanFineCounts;
    p.fine_sz.overflow = p.d_words + kPlanOverflow;
    t->d_sets = (SetParams *)ar.take(sizeof(SetParams) * nsets);
    if (p.fine) t->knn_tree->d_sets = (SetParams *)ar.take(sizeof(SetParams) * nsets);
    ctx->sync();
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    OTG_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    bool short_arena = false;
    try {
      build_plan_body<D>(p, ar);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(ctx->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      ctx->profiling = prof;
      short_arena = ar.exhausted;
      if (!short_arena) throw;
    }
    if (short_arena) {
      ctx->free(p.mem);
      p.mem = nullptr;
      continue;
    }
    ctx->profiling = prof;
    OTG_HIP_CHECK(hipStreamEndCapture(ctx->stream, &p.graph));
    break;
  }
  OTG_HIP_CHECK(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
}

BuildPlan *build_plan_create(Ctx *ctx, int dim, const double *geom, const uint64_t *offsets, uint32_t nsets,
                             const double *boxes, const uint32_t *depths, const uint32_t *caps) {
  if (!boxes) throw Error{OTG_ERR_INVALID_ARG, "build plans take every set's box"};
  if (dim < 2 || dim > 6) throw Error{OTG_ERR_INVALID_ARG, "dimension must be 2..6"};
  if (offsets[nsets] == 0) throw Error{OTG_ERR_INVALID_ARG, "plan of an empty forest"};
  std::unique_ptr<BuildPlan> p(new BuildPlan());
  p->ctx = ctx;
  p->geom = geom;
  p->n = offsets[nsets];
  p->dim = dim;
  p->t.reset(new ::otg_tree());
  p->t->ctx = ctx;
  p->t->kind = OTG_KIND_POINT;
  p->t->dim = dim;
  // a plain build of the current contents fixes the plan's shape and capacities
  Tree eager;
  eager.ctx = ctx;
  eager.kind = OTG_KIND_POINT;
  eager.dim = dim;
  build_forest(&eager, geom, OTG_MEM_DEVICE, offsets, nsets, boxes, depths, caps);
  switch (dim) {
#define OTG_CASE(DD)                                                             \
  case DD:                                                                       \
    build_plan_capture<DD>(*p, eager, offsets, nsets, boxes, depths, caps);     \
    break;
    OTG_CASE(2)
    OTG_CASE(3)
    OTG_CASE(4)
    OTG_CASE(5)
    OTG_CASE(6)
#undef OTG_CASE
  }
  publish_views(p->t.get());
  build_plan_run(p.get());
  return p.release();
}

void build_plan_run(BuildPlan *p) {
  Ctx *ctx = p->ctx;
  OTG_HIP_CHECK(hipGraphLaunch(p->exec, ctx->stream));
  ctx->sync();
  build_plan_finish(p);
}

// After a replay (alone or inside a larger graph) has completed: the host side of the
// tree follows the device.
void build_plan_finish(BuildPlan *p) {
  Tree *t = p->t.get();
  const uint32_t nsets = (uint32_t)t->sets.size();
  const uint32_t *w = p->h_words;
  std::memcpy(t->sets.data(), p->h_sets, sizeof(SetParams) * nsets);
  t->node_count = w[kPlanRefCounts + 3];
  const int D = p->dim;
  t->device_bytes = p->n * (D * 8 + 4 + 8) + t->node_count * (sizeof(NodeInfo) + 4 + 1 + 8 + 2 * D * 8);
  if (Tree *aux = t->knn_tree) {
    std::memcpy(aux->sets.data(), p->h_sets + nsets, sizeof(SetParams) * nsets);
    aux->node_count = w[kPlanFineCounts + 3];
    aux->device_bytes = p->n * 8 + aux->node_count * (sizeof(NodeInfo) + 4 + 1 + 8 + 2 * D * 8);
    t->device_bytes += aux->device_bytes;
  }
  // the reference-order storage of the last contents is stale
  std::lock_guard<std::mutex> g(t->ref_mu);
  t->ctx->free(t->d_ids_ref);
  t->ctx->free(t->d_geom_ref);
  t->ctx->free(t->d_view_ref);
  t->d_ids_ref = nullptr;
  t->d_geom_ref = nullptr;
  t->d_view_ref = nullptr;
  // a table over its capacity was left empty (no set has nodes): the host side says so too
  if (w[kPlanErr] & kErrOutOfSpace) throw Error{OTG_ERR_OUT_OF_SPACE, "point outside the given box"};
  if (w[kPlanOverflow])
    throw Error{OTG_ERR_CAPACITY, "the new contents need more nodes than the plan holds (make a new plan)"};
}

const Tree *build_plan_tree(const BuildPlan *p) { return p->t.get(); }
hipGraph_t build_plan_graph(const BuildPlan *p) { return p->graph; }
hipGraphExec_t build_plan_exec(const BuildPlan *p) { return p->exec; }
void build_plan_free(BuildPlan *p) { delete p; }

void build_forest(Tree *t, const double *geom, int mem, const uint64_t *offsets, uint32_t nsets, const double *boxes,
                  const uint32_t *depths, const uint32_t *caps) {
  const bool box = t->kind == OTG_KIND_BOX;
  switch (t->dim) {
#define OTG_CASE(DD)                                                                    \
  case DD:                                                                              \
    if (box) build_impl<DD, true>(t, geom, mem, offsets, nsets, boxes, depths, caps);  \
    else build_impl<DD, false>(t, geom, mem, offsets, nsets, boxes, depths, caps);     \
    break;
    OTG_CASE(2)
    OTG_CASE(3)
    OTG_CASE(4)
    OTG_CASE(5)
    OTG_CASE(6)
#undef OTG_CASE
    default:
      throw Error{OTG_ERR_INVALID_ARG, "dimension must be 2..6"};
  }
  publish_views(t);
}

}  // namespace otg
