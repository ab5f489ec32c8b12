// otg_query.hip -- batched queries over resident static trees.
//
// One lane per query with a short private stack (depth-first, children pushed
// so that the nearest one is popped first where the query has a distance
// order).  Every predicate that decides an id is the reference's FP64
// expression with the same operation order (built with -ffp-contract=off):
//   kNN        ot_query.h:1208-1270, AddEntityDistance :1058-1164,
//              GetBoxWallDistanceAD internal_geometry_module.h:424-458
//   range      ot_query.h:476-583
//   ray first  ot_query.h:2035-2122, RayHitTester igm:536-723
//   collision  ot_query.h:1901-1906 (pairs with GetBoxRelation == Overlapped)
// The result sets equal the reference's; only traversal order differs, which
// the reference leaves unobservable except for exact ties at the final kNN /
// ray bound (counted as measure-zero divergences by the tests).
#include <algorithm>
#include <cfloat>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "otg_device.hpp"
#include "otg_internal.hpp"

namespace otg {

constexpr int kQB = 64;  // lanes per block for traversals (1 wavefront)


// Context whose stream and allocator run a query: the calling thread's
// override (serving workers each own one) or the tree's context.
thread_local Ctx *tl_run_ctx = nullptr;
Ctx *run_ctx(const Tree *t) { return tl_run_ctx ? tl_run_ctx : t->ctx; }
void set_thread_run_ctx(Ctx *c) { tl_run_ctx = c; }

TreeView view_of(const Tree *t) {
  return TreeView{t->d_info, t->d_rend, t->d_nmin, t->d_nsize, t->d_geom, t->n, t->d_ids, t->d_sets};
}

void publish_views(Tree *t) {
  for (Tree *x = t; x; x = x->knn_tree) {
    if (!x->d_view) x->d_view = (TreeView *)x->ctx->alloc(sizeof(TreeView));
    const TreeView hv = view_of(x);
    OTG_HIP_CHECK(hipMemcpyAsync(x->d_view, &hv, sizeof(TreeView), hipMemcpyHostToDevice, x->ctx->stream));
    OTG_HIP_CHECK(hipStreamSynchronize(x->ctx->stream));
  }
}

// Runs f(tree view) for this lane's tree.  The view index is made wave-uniform
// (one tree per pass over the lanes of a wavefront), so the tree's pointers are
// loaded once into scalar registers; queries are ordered by tree, so a wavefront
// almost always finishes in one pass.
template <typename F>
__device__ __forceinline__ void with_tree(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                          uint64_t qi, F &&f) {
  const uint32_t mine = qview ? qview[qi] : 0u;
  for (;;) {
    const uint32_t v = __builtin_amdgcn_readfirstlane(mine);
    if (mine == v) {
      const TreeView tv = views[v];
      f(tv);
      return;
    }
  }
}

template <int D>
struct StackCap {
  // root + maxDepth * (2^D - 1) pending siblings
  static constexpr int kDepth = D == 2 ? 15 : (D == 3 ? 19 : (D == 4 ? 15 : (D == 5 ? 12 : 10)));
  static constexpr int value = 1 + kDepth * ((1 << D) - 1);
};

// Tree arrays live in device global memory: reading them through a global-space
// pointer gives global (or, at wave-uniform addresses, scalar) loads instead of flat
// ones, which would also hold up every LDS access behind them.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gp(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}

// A record read through a global-space pointer, 32 bits at a time.
template <typename T>
__device__ __forceinline__ T gload(const T *p) {
  static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "records of 32-bit words");
  uint32_t w[sizeof(T) / 4];
#pragma unroll
  for (size_t i = 0; i < sizeof(T) / 4; ++i) w[i] = gp(reinterpret_cast<const uint32_t *>(p))[i];
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

// ot_query.h:1033-1036 GetValueWithToleranceUpper
__device__ __forceinline__ double tol_upper(double v, double tol) { return fmax(tol, v * (1.0 + tol)); }

// A bound T >= (next double above F)^2, exact real arithmetic.  sqrt is
// correctly rounded and monotone, so d2 >= T implies sqrt(d2) > F: such
// candidates are rejected without taking the square root, and every other
// decision is made on sqrt(d2) exactly as the reference makes it.
__device__ __forceinline__ double sq_reject_bound(double F) {
  // F >= 0 and finite here; the next double above it is one unit in the last place up
  const double nf = __longlong_as_double(__double_as_longlong(F) + 1);
  return (nf * nf) * (1.0 + 0x1p-50);
}

// GetBoxWallDistanceAD(.., true) as a squared distance (0 inside); the node is
// pruned exactly when wall_distance() >= F, decided on sqrt when s2 < T.
template <int D>
__device__ __forceinline__ double wall_distance2(const double (&q)[D], const double *mn, const double *sz, bool *inside) {
  double s2 = 0.0;
  bool in = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double hs = gp(sz)[d] * 0.5;
    const double cd = fabs(gp(mn)[d] + hs - q[d]);
    in &= cd <= hs;
    const double v = smax(0.0, cd - hs);
    s2 += v * v;
  }
  *inside = in;
  return in ? 0.0 : s2;
}

// ---------------------------------------------------------------------------
// kNN (points)
// ---------------------------------------------------------------------------
// Per-query candidate list (bd/bid arrays + KnnState) with the reference's acceptance rule
// (AddEntityDistance ot_query.h:1058-1164): an entity enters when its distance
// is below F; once k entries exist F = tol_upper(k-th smallest) (assigned
// unconditionally the first time, :1124-1129), afterwards only a smaller limit
// replaces F and entries above it leave (:1138-1157).  The list keeps the KCAP
// smallest entries; a dropped entry that would still be <= F at the end marks
// the query as overflowed (reported, never silently truncated).
template <int KCAP>
struct KnnState {
  uint32_t cnt;
  double F, T, dropped;
  bool heapified, any_dropped;
};

// Compare-exchange of (distance, id) list entries: ascending distance, ties by id.
__device__ __forceinline__ void cas_id(double &da, uint32_t &ia, double &db, uint32_t &ib) {
  const bool sw = db < da || (db == da && ib < ia);
  const double tw = da;
  const uint32_t ti = ia;
  da = sw ? db : da;
  ia = sw ? ib : ia;
  db = sw ? tw : db;
  ib = sw ? ti : ib;
}

// Entries of equal distance sit in arrival order in the list; the output orders them
// by id.  The list is sorted by distance, so compare-exchanges of neighbours by
// (distance, id) only reorder equal runs, and KCAP rounds of odd-even transposition
// sort them completely.  Ties are rare: the rounds run only when one exists.
template <int KCAP>
__device__ __forceinline__ void knn_tie_order(double (&bd)[KCAP], uint32_t (&bid)[KCAP], uint32_t cnt) {
  bool tie = false;
#pragma unroll
  for (int t = 0; t + 1 < KCAP; ++t) tie |= (uint32_t)(t + 1) < cnt && bd[t] == bd[t + 1];
  if (!tie) return;
  for (int r = 0; r < KCAP; ++r) {
#pragma unroll
    for (int t = 0; t + 1 < KCAP; ++t) {
      if ((t & 1) != (r & 1)) continue;
      if ((uint32_t)(t + 1) < cnt && bd[t] == bd[t + 1] && bid[t + 1] < bid[t]) {
        const uint32_t x = bid[t];
        bid[t] = bid[t + 1];
        bid[t + 1] = x;
      }
    }
  }
}

// Batcher's odd-even merge sort on N = 2^m inputs (Knuth, TAOCP 3, 5.3.4, Algorithm M)
// as a compile-time list of compare-exchange pairs: 19 for N = 8, 63 for 16, 191 for
// 32, 543 for 64.  Fully unrolled, every index is a constant (the list stays in
// registers).
template <int N>
struct MergeNet {
  static constexpr int kCap = N * 10;
  int n = 0;
  int a[kCap] = {};
  int b[kCap] = {};
  constexpr MergeNet() {
    for (int p = 1; p < N; p <<= 1)
      for (int k = p; k >= 1; k >>= 1)
        for (int j = k % p; j + k < N; j += 2 * k)
          for (int i = 0; i < k && i + j + k < N; ++i)
            if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
              a[n] = i + j;
              b[n] = i + j + k;
              ++n;
            }
  }
};

// Where a walk leaves its list: `ids`, KCAP slots per query; or, when `direct` is set
// (a batch with one k for every query) and the list holds exactly k ids, straight into
// the rows of the batch's CSR (row qi at qi * k), so a batch in which every query has k
// ids needs no compaction.  A list of another length goes to `ids` and sets
// *irregular.
struct KnnOut {
  uint32_t *ids;
  uint32_t *direct;
  uint32_t *irregular;
};

template <int KCAP>
__device__ __forceinline__ void knn_write(const KnnOut &o, uint64_t qi, const double (&bd)[KCAP],
                                          const uint32_t (&bid)[KCAP], uint32_t cnt, uint32_t k) {
  if (o.direct && cnt == k) {
#pragma unroll
    for (int t = 0; t < KCAP; ++t)
      if ((uint32_t)t < k) o.direct[qi * k + t] = bid[t];
    return;
  }
  if (o.direct) *o.irregular = 1u;
#pragma unroll
  for (int t = 0; t < KCAP; ++t)
    if ((uint32_t)t < cnt) o.ids[qi * KCAP + t] = bid[t];
}

template <int KCAP>
__device__ __forceinline__ void knn_init(KnnState<KCAP> &S, double F0) {
  S.cnt = 0;
  S.F = F0;
  S.T = sq_reject_bound(F0);
  S.dropped = 0.0;
  S.heapified = false;
  S.any_dropped = false;
}

// GetNodeWallDistance >= F: the node would be skipped by the reference
template <int KCAP>
__device__ __forceinline__ bool knn_node_out(const KnnState<KCAP> &S, double w2, bool inside) {
  return inside ? (0.0 >= S.F) : (w2 >= S.T || sqrt(w2) >= S.F);
}

// Inserts (dist, id) into the optimistic list (ascending distance; equal distances in
// arrival order, put in id order by knn_tie_order at the output), keeping the KCAP
// smallest and remembering the smallest dropped distance.  Which of several entries
// of equal distance leaves a full list changes nothing: the dropped distance is the
// same, and a dropped distance <= F sends the query to the band pass anyway.
template <int KCAP>
__device__ __forceinline__ void knn_insert(double (&bd)[KCAP], uint32_t (&bid)[KCAP], KnnState<KCAP> &S, double dist,
                                           uint32_t id) {
  // before[t]: entry t stays ahead of the new one.  The list is sorted, so these form
  // a prefix; the new entry lands at the first t that is not before, and the entries
  // from there on move up one place: one compare and one select per slot.
  bool before[KCAP];
#pragma unroll
  for (int t = 0; t < KCAP; ++t) before[t] = (uint32_t)t < S.cnt && bd[t] <= dist;
  if (S.cnt == KCAP) {  // the largest of the KCAP + 1 entries leaves the list
    const double out_d = before[KCAP - 1] ? dist : bd[KCAP - 1];
    S.dropped = S.any_dropped ? smin(S.dropped, out_d) : out_d;
    S.any_dropped = true;
  }
#pragma unroll
  for (int t = KCAP - 1; t >= 1; --t) {
    if (!before[t]) {
      bd[t] = before[t - 1] ? dist : bd[t - 1];
      bid[t] = before[t - 1] ? id : bid[t - 1];
    }
  }
  if (!before[0]) {
    bd[0] = dist;
    bid[0] = id;
  }
  if (S.cnt < KCAP) ++S.cnt;
}

// F from the pessimistic k-th distance dk (:1120-1157): assigned at the first
// heapify, afterwards only a smaller limit replaces it and optimistic entries
// above it leave.
template <int KCAP>
__device__ __forceinline__ void knn_limit(const double (&bd)[KCAP], KnnState<KCAP> &S, double dk, double tol) {
  const double lim = tol_upper(dk, tol);
  if (!S.heapified) {
    S.heapified = true;
    S.F = lim;  // the make_heap branch assigns unconditionally (:1124-1129)
    S.T = sq_reject_bound(S.F);
  } else if (lim < S.F) {
    S.F = lim;
    S.T = sq_reject_bound(S.F);
    uint32_t keep = 0;
#pragma unroll
    for (int t = 0; t < KCAP; ++t)
      if ((uint32_t)t < S.cnt && !(bd[t] > S.F)) ++keep;
    S.cnt = keep;
  }
}

// k = KCAP - 1 (the largest k of the list's class): the k-th entry and the one slot
// that can leave sit at fixed places.  Entries 0..k-1 are <= dk <= F = tol_upper(dk),
// so only the last slot is tested against a shrunk F.
template <int KCAP>
__device__ __forceinline__ void knn_limit_top(const double (&bd)[KCAP], KnnState<KCAP> &S, double tol) {
  const double lim = tol_upper(bd[KCAP - 2], tol);
  if (!S.heapified) {
    S.heapified = true;
    S.F = lim;
    S.T = sq_reject_bound(S.F);
  } else if (lim < S.F) {
    S.F = lim;
    S.T = sq_reject_bound(S.F);
    S.cnt = (S.cnt == (uint32_t)KCAP && !(bd[KCAP - 1] > S.F)) ? KCAP : KCAP - 1;
  }
}

// Points: the optimistic and pessimistic distances coincide, so the pessimistic
// k-th is the optimistic list's k-th entry.
template <int KCAP>
__device__ __forceinline__ void knn_offer(double (&bd)[KCAP], uint32_t (&bid)[KCAP], KnnState<KCAP> &S, double dist,
                                          uint32_t id, uint32_t k, double tol) {
  knn_insert(bd, bid, S, dist, id);
  if (S.cnt >= k) {
    if (k == (uint32_t)KCAP - 1) {
      knn_limit_top(bd, S, tol);
      return;
    }
    double dk = 0.0;
#pragma unroll
    for (int t = 0; t < KCAP; ++t)
      if ((uint32_t)t == k - 1) dk = bd[t];
    knn_limit(bd, S, dk, tol);
  }
}

// The first k = KCAP - 1 entries of an empty list (k at the top of its class) from
// the first k entities of a leaf chunk: their distances go straight into the slots,
// one sorting network orders them, and F follows as after the k-th insertion.  This
// equals k sequential offers: each of them enters (checked below), an empty list
// drops nothing, and the network's (distance, id) order is the list's.  Returns
// false, the list still empty (its count is 0), when one of them would not enter.
template <int D, int KCAP>
__device__ __forceinline__ bool knn_fill(const TreeView &tv, const double (&q)[D], uint32_t e0, double (&bd)[KCAP],
                                         uint32_t (&bid)[KCAP], KnnState<KCAP> &L, double tol) {
  bool ok = true;
#pragma unroll
  for (int t = 0; t < KCAP - 1; ++t) {
    double d2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double v = q[d] - gp(tv.geom)[(uint64_t)d * tv.n + e0 + t];
      d2 += v * v;
    }
    const double dist = sqrt(d2);
    ok &= d2 < L.T && dist < L.F;
    bd[t] = dist;
    bid[t] = gp(tv.ids)[e0 + t];
  }
  if (!ok) return false;
  constexpr MergeNet<KCAP - 1> net{};
#pragma unroll
  for (int p = 0; p < net.n; ++p) cas_id(bd[net.a[p]], bid[net.a[p]], bd[net.b[p]], bid[net.b[p]]);
  L.cnt = KCAP - 1;
  knn_limit_top(bd, L, tol);
  return true;
}

// Boxes: optimistic distance into the list, pessimistic distance into the k
// smallest pessimistic values (push_heap + pop_heap of :1131-1134), F from their
// largest.
template <int KCAP>
__device__ __forceinline__ void knn_offer_box(double (&bd)[KCAP], uint32_t (&bid)[KCAP], double (&pes)[KCAP],
                                              uint32_t &npes, KnnState<KCAP> &S, double dmin, double dminmax,
                                              uint32_t id, uint32_t k, double tol) {
  knn_insert(bd, bid, S, dmin, id);
  uint32_t pos = 0;
#pragma unroll
  for (int t = 0; t < KCAP; ++t)
    if ((uint32_t)t < npes && pes[t] <= dminmax) ++pos;
  if (pos < k) {
#pragma unroll
    for (int t = KCAP - 1; t >= 1; --t)
      if ((uint32_t)t > pos && (uint32_t)t <= npes && (uint32_t)t < k) pes[t] = pes[t - 1];
#pragma unroll
    for (int t = 0; t < KCAP; ++t)
      if ((uint32_t)t == pos) pes[t] = dminmax;
    if (npes < k) ++npes;
  }
  if (npes >= k) {
    double dk = 0.0;
#pragma unroll
    for (int t = 0; t < KCAP; ++t)
      if ((uint32_t)t == k - 1) dk = pes[t];
    knn_limit(bd, S, dk, tol);
  }
}

// MinMaxDistance2 (adapters/general.h:159-210) in its operation order: squared
// optimistic (0 inside the box within tol per axis) and pessimistic distances.
template <int D>
__device__ __forceinline__ void box_minmax2(const double (&q)[D], const TreeView &tv, uint32_t j, double tol,
                                            double *min2_out, double *mm2_out) {
  double min2 = 0.0, mm2 = 0.0, far_in = DBL_MAX, largest = 0.0;
  bool inside = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double minDist = q[d] - gp(tv.geom)[(uint64_t)d * tv.n + j];
    const double maxDist = gp(tv.geom)[(uint64_t)(D + d) * tv.n + j] - q[d];
    const bool in_c = (-tol <= minDist && -tol <= maxDist);
    inside &= in_c;
    double a = minDist * minDist, b = maxDist * maxDist;
    if (b < a) {
      const double t = a;
      a = b;
      b = t;
    }
    if (inside) far_in = (b < far_in) ? b : far_in;
    if (!in_c) min2 += a;
    largest = (largest < b - a) ? b - a : largest;
    mm2 += b;
  }
  *min2_out = inside ? 0.0 : min2;
  *mm2_out = inside ? far_in : mm2 - largest;
}

// Box entities [j, je): MinMaxDistance per entity (ot_query.h:1079-1082), an
// entity enters when its optimistic distance is below F.
template <int D, int KCAP>
__device__ __forceinline__ void knn_scan_box(const TreeView &tv, const double (&q)[D], uint32_t j, uint32_t je,
                                             double (&bd)[KCAP], uint32_t (&bid)[KCAP], double (&pes)[KCAP],
                                             uint32_t &npes, KnnState<KCAP> &L, uint32_t k, double tol) {
  for (; j < je; ++j) {
    double min2, mm2;
    box_minmax2<D>(q, tv, j, tol, &min2, &mm2);
    if (min2 >= L.T) continue;
    const double dmin = sqrt(min2);
    if (dmin >= L.F) continue;
    knn_offer_box(bd, bid, pes, npes, L, dmin, sqrt(mm2), gp(tv.ids)[j], k, tol);
  }
}

// Traversal stack for the generic walk: entries (node, squared wall distance at
// push time, as a float rounded down) in private memory with the worst-case depth
// (root + maxDepth x (2^D - 1) pending siblings).  An entry the rounding leaves
// unpruned only costs its distance tests.
template <int D>
struct ArrayStack {
  static constexpr int cap = StackCap<D>::value;
  uint32_t n[cap];
  float w[cap];  // rounded down: a pop prunes only what the exact value prunes
  __device__ __forceinline__ uint32_t node(int i) const { return n[i]; }
  __device__ __forceinline__ double wall(int i) const { return (double)w[i]; }
  __device__ __forceinline__ void set(int i, uint32_t nd, double wl) {
    n[i] = nd;
    w[i] = __double2float_rd(wl);
  }
};

// Level entries per lane: only internal nodes push one, and they sit at depths 0..maxDepth-1 on
// the current path, so kMaxDepth entries always suffice (19 x 8 B x 128 lanes = 19 KiB per
// block).
constexpr int kLdsStack = kMaxDepth;


// One lane per query (queries sorted by (set, Morton code) so neighbouring lanes
// walk nearby nodes); depth-first with the nearest child popped first.  Stack
// entries carry the child's squared wall distance computed when it was pushed,
// so a pop re-checks it against the current F without reloading the node box.
// Returns false (writing nothing) when the stack would overflow.
template <int D, int KCAP, bool BOX, typename Stack>
__device__ __forceinline__ bool knn_query(const TreeView &tv, uint64_t qi, const double *__restrict__ qpts,
                                          const uint32_t *__restrict__ qset, uint32_t k, double F0, double tol,
                                          Stack &S, KnnOut out_ids, uint32_t *__restrict__ out_cnt,
                                          uint32_t *__restrict__ overflow, uint32_t *__restrict__ band,
                                          double *__restrict__ out_F) {
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  double q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = qpts[qi * D + d];
  double bd[KCAP];
  uint32_t bid[KCAP];
  KnnState<KCAP> L;
  knn_init(L, F0);
  double pes[KCAP];
  uint32_t npes = 0;
  if (sp.node_cnt != 0 && k != 0) {
    int top = 0;
    {
      const uint32_t root = (uint32_t)sp.node_off;
      bool in;
      S.set(0, root, wall_distance2<D>(q, &tv.nmin[(uint64_t)root * D], &tv.nsize[(uint64_t)root * D], &in));
      top = 1;
    }
    while (top > 0) {
      --top;
      const uint32_t node = S.node(top);
      // wall2 is 0 inside the node box, and (0 >= F) is the reference's test there too
      if (knn_node_out(L, S.wall(top), false)) continue;
      const NodeInfo ni = gload(tv.info + node);
      if constexpr (BOX)
        knn_scan_box<D, KCAP>(tv, q, ni.ebegin, ni.ebegin + ni.ecount, bd, bid, pes, npes, L, k, tol);
      else
        knn_scan<D, KCAP>(tv, q, ni.ebegin, ni.ebegin + ni.ecount, bd, bid, L, k, tol);
      if (ni.ccount) {
        // push children still within F, farthest first so the nearest pops first
        const int base = top;
        for (uint32_t c = 0; c < ni.ccount; ++c) {
          const uint32_t ch = ni.cbegin + c;
          bool cin;
          const double wc = wall_distance2<D>(q, &tv.nmin[(uint64_t)ch * D], &tv.nsize[(uint64_t)ch * D], &cin);
          if (knn_node_out(L, wc, false)) continue;
          if (top == Stack::cap) return false;
          int p = top;
          while (p > base && S.wall(p - 1) < wc) {
            S.set(p, S.node(p - 1), S.wall(p - 1));
            --p;
          }
          S.set(p, ch, wc);
          ++top;
        }
      }
    }
    // the band below F holds more ids than the list: the query is redone by k_knn_band
    if (L.any_dropped && L.dropped <= L.F) band[atomicAdd(overflow, 1u)] = (uint32_t)qi;
  }
  out_cnt[qi] = L.cnt;
  if (out_F) out_F[qi] = L.F;
  knn_tie_order(bd, bid, L.cnt);
  knn_write(out_ids, qi, bd, bid, L.cnt, k);
  return true;
}

// Distance tests of entities [j, je) against the candidate list, in chunks of 64.
// Pass 1 computes every distance of the chunk (coordinates of four entities are
// loaded before any is tested, so a lane keeps several memory requests in flight)
// and marks the entities below the lane's bound at the start of the chunk.  Pass 2
// offers the marked entities in entity order, each re-checked against the bound at
// that moment (the bound only shrinks), so the list evolves exactly as in a plain
// sequential scan.  The costly list insertion then runs once per marked entity of
// the busiest lane instead of once per entity position that any lane accepts.
template <int D, int KCAP>
__device__ __forceinline__ void knn_scan(const TreeView &tv, const double (&q)[D], uint32_t j, uint32_t je,
                                         double (&bd)[KCAP], uint32_t (&bid)[KCAP], KnnState<KCAP> &L, uint32_t k,
                                         double tol) {
  constexpr int U = D <= 3 ? 4 : (D == 4 ? 2 : 1);  // entities loaded ahead (register budget)
  for (uint32_t base = j; base < je; base += 64) {
    const uint32_t cend = min(je, base + 64u);
    uint32_t e = base;
    // an empty list and k at the top of its class: the first k entities fill it at once
    if (L.cnt == 0 && k == (uint32_t)KCAP - 1 && cend - base >= k && knn_fill<D, KCAP>(tv, q, base, bd, bid, L, tol))
      e = base + k;
    const double T0 = L.T;
    uint64_t mark = 0;
    for (; e + U <= cend; e += U) {
      double d2[U];
#pragma unroll
      for (int u = 0; u < U; ++u) d2[u] = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = gp(tv.geom)[(uint64_t)d * tv.n + e + u];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double v = q[d] - c[u];
          d2[u] += v * v;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (d2[u] < T0) mark |= 1ull << (e + u - base);
    }
    for (; e < cend; ++e) {
      double d2 = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double v = q[d] - gp(tv.geom)[(uint64_t)d * tv.n + e];
        d2 += v * v;
      }
      if (d2 < T0) mark |= 1ull << (e - base);
    }
    // the next marked entity's coordinates and id are loaded before the current one is
    // offered, so the reload overlaps the list update
    if (!mark) continue;
    uint32_t x = base + (uint32_t)__builtin_ctzll(mark);
    mark &= mark - 1;
    double c[D];
#pragma unroll
    for (int d = 0; d < D; ++d) c[d] = gp(tv.geom)[(uint64_t)d * tv.n + x];
    uint32_t id = gp(tv.ids)[x];
    for (;;) {
      double d2 = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double v = q[d] - c[d];
        d2 += v * v;
      }
      const uint32_t cur_id = id;
      const bool more = mark != 0;
      if (more) {
        x = base + (uint32_t)__builtin_ctzll(mark);
        mark &= mark - 1;
#pragma unroll
        for (int d = 0; d < D; ++d) c[d] = gp(tv.geom)[(uint64_t)d * tv.n + x];
        id = gp(tv.ids)[x];
      }
      if (d2 < L.T) {
        const double dist = sqrt(d2);
        if (dist < L.F) knn_offer(bd, bid, L, dist, cur_id, k, tol);
      }
      if (!more) break;
    }
  }
}

// Compare-exchange on (key, index) pairs held in registers (static indices).
__device__ __forceinline__ void cas(double &wa, uint32_t &ia, double &wb, uint32_t &ib) {
  if (wb < wa) {
    const double tw = wa;
    wa = wb;
    wb = tw;
    const uint32_t ti = ia;
    ia = ib;
    ib = ti;
  }
}

// Stack of tree levels for D <= 3: one entry per node on the current path whose
// children are still pending -- the index of its first child and the pending
// children's offsets in nearest-first order, packed D bits each, with their count
// in the top 4 bits.  The depth is bounded by the tree depth (<= 19 levels).
struct LdsLevelStack {
  static constexpr int cap = kLdsStack;
  uint32_t *cb;
  uint32_t *pk;
};

// kNN walk with level entries: a pop takes the nearest pending child, recomputes
// its wall distance and, when the current radius excludes it, drops the level
// (its remaining children are farther).  Same visits and acceptance rule as
// knn_query; only the order of equally distant siblings may differ.
template <int D, int KCAP, bool BOX>
__device__ __forceinline__ bool knn_query_levels(const TreeView &tv, uint64_t qi, const double *__restrict__ qpts,
                                                 const uint32_t *__restrict__ qset, uint32_t k, double F0, double tol,
                                                 LdsLevelStack &S, KnnOut out_ids,
                                                 uint32_t *__restrict__ out_cnt, uint32_t *__restrict__ overflow,
                                                 uint32_t *__restrict__ band, double *__restrict__ out_F) {
  static_assert(D <= 3, "packed child order needs D * 2^D <= 28 bits");
  constexpr int NC = 1 << D;
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  double q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = qpts[qi * D + d];
  double bd[KCAP];
  uint32_t bid[KCAP];
  KnnState<KCAP> L;
  knn_init(L, F0);
  double pes[KCAP];
  uint32_t npes = 0;
  if (sp.node_cnt != 0 && k != 0) {
    int top = 0;
    uint32_t node = (uint32_t)sp.node_off;
    // the node record is loaded together with the box that decides the visit
    NodeInfo ni = gload(tv.info + node);
    bool have = true;  // `node` is to be visited
    {
      bool in;
      const double w = wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in);
      have = !knn_node_out(L, w, false);
    }
    for (;;) {
      if (have) {
        if constexpr (BOX)
          knn_scan_box<D, KCAP>(tv, q, ni.ebegin, ni.ebegin + ni.ecount, bd, bid, pes, npes, L, k, tol);
        else
          knn_scan<D, KCAP>(tv, q, ni.ebegin, ni.ebegin + ni.ecount, bd, bid, L, k, tol);
        if (ni.ccount) {
          double w[NC];
          uint32_t ix[NC];
          uint32_t cnt = 0;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            ix[c] = (uint32_t)c;
            w[c] = __builtin_inf();
            if ((uint32_t)c < ni.ccount) {
              const uint32_t ch = ni.cbegin + c;
              bool cin;
              const double wc =
                  wall_distance2<D>(q, &tv.nmin[(uint64_t)ch * D], &tv.nsize[(uint64_t)ch * D], &cin);
              if (!knn_node_out(L, wc, false)) {
                w[c] = fmin(wc, DBL_MAX);
                ++cnt;
              }
            }
          }
          if (NC == 4) {
            cas(w[0], ix[0], w[1], ix[1]); cas(w[2], ix[2], w[3], ix[3]);
            cas(w[0], ix[0], w[2], ix[2]); cas(w[1], ix[1], w[3], ix[3]);
            cas(w[1], ix[1], w[2], ix[2]);
          } else {  // Batcher odd-even merge sort of 8
            cas(w[0], ix[0], w[1], ix[1]); cas(w[2], ix[2], w[3], ix[3]);
            cas(w[4], ix[4], w[5], ix[5]); cas(w[6], ix[6], w[7], ix[7]);
            cas(w[0], ix[0], w[2], ix[2]); cas(w[1], ix[1], w[3], ix[3]);
            cas(w[4], ix[4], w[6], ix[6]); cas(w[5], ix[5], w[7], ix[7]);
            cas(w[1], ix[1], w[2], ix[2]); cas(w[5], ix[5], w[6], ix[6]);
            cas(w[0], ix[0], w[4], ix[4]); cas(w[1], ix[1], w[5], ix[5]);
            cas(w[2], ix[2], w[6], ix[6]); cas(w[3], ix[3], w[7], ix[7]);
            cas(w[2], ix[2], w[4], ix[4]); cas(w[3], ix[3], w[5], ix[5]);
            cas(w[1], ix[1], w[2], ix[2]); cas(w[3], ix[3], w[4], ix[4]);
            cas(w[5], ix[5], w[6], ix[6]);
          }
          if (cnt) {
            uint32_t packed = cnt << 28;
#pragma unroll
            for (int c = 0; c < NC; ++c)
              if ((uint32_t)c < cnt) packed |= ix[c] << (D * c);
            if (top == LdsLevelStack::cap) return false;
            S.cb[top * kQB] = ni.cbegin;
            S.pk[top * kQB] = packed;
            ++top;
          }
        }
      }
      if (top == 0) break;
      // nearest pending child of the innermost level
      const uint32_t packed = S.pk[(top - 1) * kQB];
      const uint32_t cnt = packed >> 28;
      node = S.cb[(top - 1) * kQB] + (packed & (NC - 1));
      ni = gload(tv.info + node);
      if (cnt == 1) --top;
      else S.pk[(top - 1) * kQB] = ((cnt - 1) << 28) | ((packed & 0x0FFFFFFFu) >> D);
      bool in;
      const double w = wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in);
      have = !knn_node_out(L, w, false);
      if (!have && cnt > 1) --top;  // the rest of this level is farther still
    }
    // the band below F holds more ids than the list: the query is redone by k_knn_band
    if (L.any_dropped && L.dropped <= L.F) band[atomicAdd(overflow, 1u)] = (uint32_t)qi;
  }
  out_cnt[qi] = L.cnt;
  if (out_F) out_F[qi] = L.F;
  knn_tie_order(bd, bid, L.cnt);
  knn_write(out_ids, qi, bd, bid, L.cnt, k);
  return true;
}

// views/qview: the trees of this launch and each query's tree (qview null: one tree);
// Levels of a 4-D..6-D walk: one per internal node on the path, and internal nodes
// sit at depths below the tree depth, which is at most 63 / D.
template <int D>
struct MaskLevels {
  static constexpr int value = (D >= 2 && D <= 6) ? 63 / D : 1;
};

// kNN walk for 4-D..6-D with one LDS entry per level: the first child and a bit
// mask of the children within the radius when their parent was visited (2^D <= 64).
// The nearest of them is visited at once; the others are taken in index order, and
// each is tested against the radius of that moment, as the reference's queue pop
// does.  Same visits rule and acceptance rule as knn_query (only the order in
// which siblings are met differs); no per-lane stack in scratch memory.
template <int D, int KCAP, bool BOX>
__device__ __forceinline__ bool knn_query_masks(const TreeView &tv, uint64_t qi, const double *__restrict__ qpts,
                                                const uint32_t *__restrict__ qset, uint32_t k, double F0, double tol,
                                                uint32_t *s_cb, uint32_t *s_lo, uint32_t *s_hi,
                                                KnnOut out_ids, uint32_t *__restrict__ out_cnt,
                                                uint32_t *__restrict__ overflow, uint32_t *__restrict__ band,
                                                double *__restrict__ out_F) {
  static_assert(D >= 4 && D <= 6, "child masks of 2^D bits need D <= 6");
  constexpr int cap = MaskLevels<D>::value;
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  double q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = qpts[qi * D + d];
  double bd[KCAP];
  uint32_t bid[KCAP];
  KnnState<KCAP> L;
  knn_init(L, F0);
  double pes[KCAP];
  uint32_t npes = 0;
  if (sp.node_cnt != 0 && k != 0) {
    int top = 0;
    uint32_t node = (uint32_t)sp.node_off;
    bool have;
    {
      bool in;
      const double w = wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in);
      have = !knn_node_out(L, w, false);
    }
    for (;;) {
      if (have) {
        const NodeInfo ni = gload(tv.info + node);
        if constexpr (BOX)
          knn_scan_box<D, KCAP>(tv, q, ni.ebegin, ni.ebegin + ni.ecount, bd, bid, pes, npes, L, k, tol);
        else
          knn_scan<D, KCAP>(tv, q, ni.ebegin, ni.ebegin + ni.ecount, bd, bid, L, k, tol);
        if (ni.ccount) {
          uint64_t mask = 0;
          double best = __builtin_inf();
          uint32_t bi = 0;
          for (uint32_t c = 0; c < ni.ccount; ++c) {
            const uint32_t ch = ni.cbegin + c;
            bool cin;
            const double wc = wall_distance2<D>(q, &tv.nmin[(uint64_t)ch * D], &tv.nsize[(uint64_t)ch * D], &cin);
            if (knn_node_out(L, wc, false)) continue;
            mask |= 1ull << c;
            if (wc < best) {
              best = wc;
              bi = c;
            }
          }
          if (mask) {
            mask &= ~(1ull << bi);
            if (mask) {
              if (top == cap) return false;
              s_cb[top * kQB] = ni.cbegin;
              s_lo[top * kQB] = (uint32_t)mask;
              s_hi[top * kQB] = (uint32_t)(mask >> 32);
              ++top;
            }
            node = ni.cbegin + bi;  // within the radius an instant ago: visited next
            continue;
          }
        }
      }
      if (top == 0) break;
      // next pending child of the innermost level, re-tested against the radius now
      uint64_t m = ((uint64_t)s_hi[(top - 1) * kQB] << 32) | s_lo[(top - 1) * kQB];
      node = s_cb[(top - 1) * kQB] + (uint32_t)__builtin_ctzll(m);
      m &= m - 1;
      if (m) {
        s_lo[(top - 1) * kQB] = (uint32_t)m;
        s_hi[(top - 1) * kQB] = (uint32_t)(m >> 32);
      } else {
        --top;
      }
      bool in;
      const double w = wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in);
      have = !knn_node_out(L, w, false);
    }
    // the band below F holds more ids than the list: the query is redone by k_knn_band
    if (L.any_dropped && L.dropped <= L.F) band[atomicAdd(overflow, 1u)] = (uint32_t)qi;
  }
  out_cnt[qi] = L.cnt;
  if (out_F) out_F[qi] = L.F;
  knn_tie_order(bd, bid, L.cnt);
  knn_write(out_ids, qi, bd, bid, L.cnt, k);
  return true;
}

// Next 64-query tile of a persistent launch: one atomic per wavefront, broadcast.
__device__ __forceinline__ uint64_t next_tile(uint32_t *tiles) {
  uint32_t t = 0;
  if ((threadIdx.x & 63) == 0) t = atomicAdd(tiles, 1u);
  return (uint64_t)__shfl(t, 0, 64);
}

// One query of a kNN launch (one lane): its tree, then the walk of its dimension.
template <int D, int KCAP, bool BOX>
__device__ __forceinline__ void knn_lane(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                         const double *__restrict__ qpts, const uint32_t *__restrict__ qset,
                                         const uint32_t *__restrict__ qk, const uint32_t *__restrict__ order,
                                         uint64_t ti, uint32_t k, double F0, double tol,
                                         KnnOut out_ids, uint32_t *__restrict__ out_cnt,
                                         uint32_t *__restrict__ overflow, uint32_t *__restrict__ band,
                                         double *__restrict__ out_F, uint32_t *s_n, uint32_t *s_w) {
  constexpr int ML = MaskLevels<D>::value;
  const uint64_t qi = order ? order[ti] : ti;
  const uint32_t kq = qk ? qk[qi] : k;
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    if constexpr (D <= 3) {
      LdsLevelStack S{&s_n[threadIdx.x], &s_w[threadIdx.x]};
      knn_query_levels<D, KCAP, BOX>(tv, qi, qpts, qset, kq, F0, tol, S, out_ids, out_cnt, overflow, band, out_F);
    } else if constexpr (D <= 6) {
      knn_query_masks<D, KCAP, BOX>(tv, qi, qpts, qset, kq, F0, tol, &s_n[threadIdx.x], &s_w[threadIdx.x],
                                    &s_w[ML * kQB + threadIdx.x], out_ids, out_cnt, overflow, band, out_F);
    } else {
      ArrayStack<D> S;
      knn_query<D, KCAP, BOX>(tv, qi, qpts, qset, kq, F0, tol, S, out_ids, out_cnt, overflow, band, out_F);
    }
  });
}

// 2-D/3-D: level entries in LDS, one per tree level.  4-D..6-D: one child mask per
// level in LDS.  7-D and up: the worst-case private stack.  None can overflow.
static_assert(kLdsStack >= kMaxDepth, "one level entry per internal depth");
// 2-D/3-D: level stack (first child, packed order); 4-D..6-D: first child and the two
// halves of the child mask, MaskLevels entries each
#define OTG_KNN_LDS(D)                                                        \
  __shared__ uint32_t s_n[(D <= 3 ? kLdsStack : MaskLevels<D>::value) * kQB]; \
  __shared__ uint32_t s_w[(D <= 3 ? kLdsStack : 2 * MaskLevels<D>::value) * kQB];

// One 64-query tile per block.  qk: per-query k (null: k for all).
template <int D, int KCAP, bool BOX>
__global__ void __launch_bounds__(kQB, 3) k_knn(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                              const double *__restrict__ qpts, const uint32_t *__restrict__ qset,
                                              const uint32_t *__restrict__ qk, const uint32_t *__restrict__ order,
                                              uint64_t nq, uint32_t k, double F0, double tol,
                                              KnnOut out_ids, uint32_t *__restrict__ out_cnt,
                                              uint32_t *__restrict__ overflow, uint32_t *__restrict__ band,
                                              double *__restrict__ out_F) {
  OTG_KNN_LDS(D)
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nq) return;
  knn_lane<D, KCAP, BOX>(views, qview, qpts, qset, qk, order, ti, k, F0, tol, out_ids, out_cnt, overflow, band, out_F,
                         s_n, s_w);
}

// Persistent launch of the same walk: each block takes 64-query tiles from the counter
// (zeroed before the launch) until none is left, so uneven sets finish as one pool.
// A separate kernel: the tile loop keeps every argument live, which costs the
// one-tile kernel registers.
template <int D, int KCAP, bool BOX>
__global__ void __launch_bounds__(kQB, 3)
    k_knn_persistent(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                     const double *__restrict__ qpts, const uint32_t *__restrict__ qset,
                     const uint32_t *__restrict__ qk, const uint32_t *__restrict__ order, uint64_t nq, uint32_t k,
                     double F0, double tol, KnnOut out_ids, uint32_t *__restrict__ out_cnt,
                     uint32_t *__restrict__ overflow, uint32_t *__restrict__ band, double *__restrict__ out_F,
                     uint32_t *__restrict__ tiles) {
  OTG_KNN_LDS(D)
  const uint64_t ntile = (nq + kQB - 1) / kQB;
  for (uint64_t tile = next_tile(tiles); tile < ntile; tile = next_tile(tiles)) {
    const uint64_t ti = tile * kQB + threadIdx.x;
    if (ti < nq)
      knn_lane<D, KCAP, BOX>(views, qview, qpts, qset, qk, order, ti, k, F0, tol, out_ids, out_cnt, overflow, band,
                             out_F, s_n, s_w);
  }
}
#undef OTG_KNN_LDS

// kNN of point trees with one wavefront per query (2-D..6-D; where it is launched:
// kKnnWave).  A node has up to 2^D <= 64 children, so the lanes test a node's children
// (one each) or a 64-entity chunk of a leaf together; the candidate list is spread
// over the lanes (slot t in lane t).  The walk is knn_query_masks's and every decision
// is the same, made once per query instead of once per lane: the same visit order
// (nearest child at once, the rest in index order, each re-tested when taken), a
// chunk's candidates marked against the bound at its start and offered in index order
// against the bound of that moment, the same insertion, limit and band rule
// (knn_insert, knn_limit), so the result is the masks walk's (in 2-D / 3-D the level
// walk meets equally distant siblings in another order, as the masks walk does: the
// ids are the same).  Four queries per block.
constexpr int kWaveQueries = 4;
// Used for 6-D (measured on 8M clustered points, 1M queries k = 16: 133.8 -> 120.8 ms)
// and for 2-D / 3-D batches of fewer than kWaveMaxQueries queries (the batches that
// also skip the query sort), which one lane per query leaves too few wavefronts to fill
// the GPU with.  Measured on 1M uniform 3-D points, k = 8, lane walk vs wavefront walk:
// 1k queries 0.317 vs 0.115 ms, 10k 0.419 vs 0.274, 50k 0.574 vs 0.966, 1M 2.107 vs
// 14.374.  4-D and 5-D keep the lane walk.
template <int D, int KCAP, bool BOX>
constexpr bool kKnnWave = (D == 6 || D <= 3) && !BOX && KCAP <= 64;
constexpr uint64_t kWaveMaxQueries = 16384;
template <int D>
static bool knn_wave_batch(uint64_t nq) {
  return D == 6 || nq < kWaveMaxQueries;
}

template <int D, int KCAP>
__global__ void __launch_bounds__(64 * kWaveQueries)
    k_knn_wave(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
               const double *__restrict__ qpts, const uint32_t *__restrict__ qset, const uint32_t *__restrict__ qk,
               const uint32_t *__restrict__ order, uint64_t nq, uint32_t k, double F0, double tol, KnnOut out,
               uint32_t *__restrict__ out_cnt, uint32_t *__restrict__ overflow, uint32_t *__restrict__ band,
               double *__restrict__ out_F) {
  static_assert(D >= 2 && D <= 6 && KCAP <= 64, "one child per lane and one list slot per lane");
  constexpr int ML = MaskLevels<D>::value;
  __shared__ uint32_t s_cb[kWaveQueries][ML], s_lo[kWaveQueries][ML], s_hi[kWaveQueries][ML];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // the query, its tree: scalar registers
  const uint64_t ti = (uint64_t)blockIdx.x * kWaveQueries + wv;
  if (ti >= nq) return;  // wavefront-uniform
  const uint64_t qi = order ? order[ti] : ti;
  const uint32_t kq = qk ? qk[qi] : k;
  const TreeView tv = views[qview ? qview[qi] : 0u];
  const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
  double q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = qpts[qi * D + d];
  // the list: slot `lane` (lanes >= KCAP hold nothing); the rest is the same in every lane
  double bd = 0.0;
  uint32_t bid = 0;
  uint32_t cnt = 0;
  double F = F0, T = sq_reject_bound(F0), dropped = 0.0;
  bool heapified = false, any_dropped = false;
  auto node_out = [&](double w2) { return w2 >= T || sqrt(w2) >= F; };
  // knn_insert + knn_limit on the spread list
  auto offer = [&](double dist, uint32_t id) {
    const uint32_t pos = (uint32_t)__popcll(__ballot(lane < cnt && bd <= dist));
    if (cnt == (uint32_t)KCAP) {
      const double last = __shfl(bd, KCAP - 1, 64);
      const double out_d = pos == (uint32_t)KCAP ? dist : last;
      dropped = any_dropped ? smin(dropped, out_d) : out_d;
      any_dropped = true;
    }
    const double up_d = __shfl_up(bd, 1, 64);
    const uint32_t up_i = __shfl_up(bid, 1, 64);
    if (lane == pos) {
      bd = dist;
      bid = id;
    } else if (lane > pos) {
      bd = up_d;
      bid = up_i;
    }
    if (cnt < (uint32_t)KCAP) ++cnt;
    if (cnt >= kq) {
      const double lim = tol_upper(__shfl(bd, (int)kq - 1, 64), tol);
      if (!heapified) {
        heapified = true;
        F = lim;
        T = sq_reject_bound(F);
      } else if (lim < F) {
        F = lim;
        T = sq_reject_bound(F);
        cnt = (uint32_t)__popcll(__ballot(lane < cnt && !(bd > F)));
      }
    }
  };
  // knn_scan: chunks of 64 entities, one per lane
  auto scan = [&](uint32_t j, uint32_t je) {
    for (uint32_t base = j; base < je; base += 64) {
      const uint32_t e = base + lane;
      const double T0 = T;
      double d2 = 0.0;
      const bool valid = e < je;
      uint32_t eid = 0;  // loaded with the coordinates: an offer takes it from its lane
      if (valid) {
        eid = gp(tv.ids)[e];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const double v = q[d] - gp(tv.geom)[(uint64_t)d * tv.n + e];
          d2 += v * v;
        }
      }
      uint64_t mark = __ballot(valid && d2 < T0);
      while (mark) {
        const uint32_t x = (uint32_t)__builtin_ctzll(mark);
        mark &= mark - 1;
        const double dx2 = __shfl(d2, (int)x, 64);
        const uint32_t xid = __shfl(eid, (int)x, 64);
        if (dx2 < T) {
          const double dist = sqrt(dx2);
          if (dist < F) offer(dist, xid);
        }
      }
    }
  };
  if (sp.node_cnt != 0 && kq != 0) {
    int top = 0;
    uint32_t node = (uint32_t)sp.node_off;
    bool in;
    // a node's record is loaded with the box that decides its visit (children: one per lane)
    NodeInfo ni = gload(tv.info + node);
    bool have = !node_out(wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in));
    for (;;) {
      if (have) {
        scan(ni.ebegin, ni.ebegin + ni.ecount);
        if (ni.ccount) {
          // child `lane`: within the radius?  The nearest of them (lowest index on ties)
          double wc = __builtin_inf();
          bool keep = false;
          NodeInfo ci{};
          if (lane < ni.ccount) {
            const uint32_t ch = ni.cbegin + lane;
            ci = gload(tv.info + ch);
            bool cin;
            wc = wall_distance2<D>(q, &tv.nmin[(uint64_t)ch * D], &tv.nsize[(uint64_t)ch * D], &cin);
            keep = !node_out(wc);
          }
          uint64_t mask = __ballot(keep);
          if (mask) {
            double bw = keep ? wc : __builtin_inf();
            uint32_t bi = keep && wc < __builtin_inf() ? lane : 64u;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
              const double ow = __shfl_xor(bw, o, 64);
              const uint32_t oi = __shfl_xor(bi, o, 64);
              if (ow < bw || (ow == bw && oi < bi)) {
                bw = ow;
                bi = oi;
              }
            }
            if (bi == 64u) bi = 0;  // no child nearer than infinity: knn_query_masks's first index
            mask &= ~(1ull << bi);
            if (mask) {
              if (top == ML) break;  // cannot happen: one level per internal depth
              if (lane == 0) {
                s_cb[wv][top] = ni.cbegin;
                s_lo[wv][top] = (uint32_t)mask;
                s_hi[wv][top] = (uint32_t)(mask >> 32);
              }
              ++top;
            }
            node = ni.cbegin + bi;  // within the radius an instant ago: visited next
            ni.ebegin = __shfl(ci.ebegin, (int)bi, 64);
            ni.ecount = __shfl(ci.ecount, (int)bi, 64);
            ni.cbegin = __shfl(ci.cbegin, (int)bi, 64);
            ni.ccount = __shfl(ci.ccount, (int)bi, 64);
            continue;
          }
        }
      }
      if (top == 0) break;
      // next pending child of the innermost level, re-tested against the radius now
      __builtin_amdgcn_wave_barrier();
      uint64_t m = ((uint64_t)s_hi[wv][top - 1] << 32) | s_lo[wv][top - 1];
      node = s_cb[wv][top - 1] + (uint32_t)__builtin_ctzll(m);
      m &= m - 1;
      __builtin_amdgcn_wave_barrier();
      if (m) {
        if (lane == 0) {
          s_lo[wv][top - 1] = (uint32_t)m;
          s_hi[wv][top - 1] = (uint32_t)(m >> 32);
        }
      } else {
        --top;
      }
      ni = gload(tv.info + node);
      have = !node_out(wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in));
    }
    if (lane == 0 && any_dropped && dropped <= F) band[atomicAdd(overflow, 1u)] = (uint32_t)qi;
  }
  if (lane == 0) {
    out_cnt[qi] = cnt;
    if (out_F) out_F[qi] = F;
  }
  // knn_tie_order: equal distances in id order, by odd-even transposition over the lanes
  const bool tie = __ballot(lane + 1 < cnt && bd == __shfl_down(bd, 1, 64)) != 0ull;
  if (tie) {
    for (int r = 0; r < KCAP; ++r) {
      const bool left = (lane & 1u) == (uint32_t)(r & 1);
      const int partner = left ? (int)lane + 1 : (int)lane - 1;
      const double pd = __shfl(bd, partner & 63, 64);
      const uint32_t pi = __shfl(bid, partner & 63, 64);
      if (left) {
        if (lane + 1 < cnt && pd == bd && pi < bid) bid = pi;
      } else if (partner >= 0 && lane < cnt && pd == bd && bid < pi) {
        bid = pi;
      }
    }
  }
  if (out.direct && cnt == kq) {
    if (lane < kq) out.direct[qi * kq + lane] = bid;
  } else {
    if (out.direct && lane == 0) *out.irregular = 1u;
    if (lane < cnt) out.ids[qi * KCAP + lane] = bid;
  }
}

// Queries whose band below the final F held more ids than the register list: a
// fixed-radius walk collects every entity with (optimistic) distance below F --
// the same arithmetic as the list pass -- counted first, then written as
// (distance, id) pairs and sorted per query.
template <int D, bool BOX, bool FILL>
__global__ void __launch_bounds__(kQB) k_knn_band(const TreeView *__restrict__ views,
                                                   const uint32_t *__restrict__ qview, const double *__restrict__ qpts,
                                                   const uint32_t *__restrict__ qset, const uint32_t *__restrict__ band,
                                                   uint32_t nband, const double *__restrict__ Fq, double tol,
                                                   uint32_t *__restrict__ counts, const uint32_t *__restrict__ offsets,
                                                   double *__restrict__ out_d, uint32_t *__restrict__ out_id) {
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nband) return;
  const uint64_t qi = band[ti];
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
    const double F = Fq[qi], T = sq_reject_bound(F);
    double q[D];
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = qpts[qi * D + d];
    uint32_t cnt = 0;
    const uint64_t o = FILL ? offsets[ti] : 0;
    if (sp.node_cnt != 0) {
      uint32_t stack[StackCap<D>::value];
      int top = 0;
      stack[top++] = (uint32_t)sp.node_off;
      while (top > 0) {
        const uint32_t node = stack[--top];
        bool in;
        const double w2 = wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in);
        if (w2 >= T || sqrt(w2) >= F) continue;
        const NodeInfo ni = gload(tv.info + node);
        for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
          double x2, mm2 = 0.0;
          if (BOX) {
            box_minmax2<D>(q, tv, j, tol, &x2, &mm2);
          } else {
            x2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
              const double v = q[d] - tv.geom[(uint64_t)d * tv.n + j];
              x2 += v * v;
            }
          }
          if (x2 >= T) continue;
          const double dist = sqrt(x2);
          if (!(dist < F)) continue;
          if (FILL) {
            out_d[o + cnt] = dist;
            out_id[o + cnt] = tv.ids[j];
          }
          ++cnt;
        }
        for (uint32_t c = 0; c < ni.ccount; ++c) stack[top++] = ni.cbegin + c;
      }
    }
    if (!FILL) counts[ti] = cnt;
  });
}

// kNN with k above the register lists (k > 64): the final radius of each query,
// F = tol_upper(k-th smallest pessimistic distance) (points: the distance) among
// the entities whose optimistic distance is below the running F, or F0 when fewer
// than k are -- the bound the reference's AddEntityDistance ends with
// (ot_query.h:1058-1164: F assigned at the k-th entry, lowered afterwards).  A
// depth-first walk keeps the k smallest pessimistic distances in a max-heap of k
// doubles in global memory (heap + ti * k); nodes and entities are skipped exactly
// where the reference skips them (wall or optimistic distance >= F).  The ids are
// then collected by the band pass with this F, which returns {optimistic < F}
// sorted by (distance, id): the reference's list.
template <int D, bool BOX>
__global__ void __launch_bounds__(kQB) k_knn_radius(const TreeView *__restrict__ views,
                                                     const uint32_t *__restrict__ qview, const double *__restrict__ qpts,
                                                     const uint32_t *__restrict__ qset, uint64_t q0, uint64_t nlist,
                                                     uint32_t k, double F0, double tol, double *__restrict__ heap,
                                                     double *__restrict__ Fq) {
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nlist) return;
  const uint64_t qi = q0 + ti;
  double *h = heap + ti * k;
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
    double F = F0, T = sq_reject_bound(F0);
    double q[D];
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = qpts[qi * D + d];
    uint32_t hn = 0;
    if (sp.node_cnt != 0) {
      uint32_t stack[StackCap<D>::value];
      int top = 0;
      stack[top++] = (uint32_t)sp.node_off;
      while (top > 0) {
        const uint32_t node = stack[--top];
        bool in;
        const double w2 = wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in);
        if (w2 >= T || sqrt(w2) >= F) continue;
        const NodeInfo ni = gload(tv.info + node);
        for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
          double x2, mm2 = 0.0;
          if (BOX) {
            box_minmax2<D>(q, tv, j, tol, &x2, &mm2);
          } else {
            x2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
              const double v = q[d] - tv.geom[(uint64_t)d * tv.n + j];
              x2 += v * v;
            }
          }
          if (x2 >= T || !(sqrt(x2) < F)) continue;
          const double pes = BOX ? sqrt(mm2) : sqrt(x2);
          if (hn < k) {
            uint32_t i = hn++;  // sift up
            while (i > 0) {
              const uint32_t par = (i - 1) / 2;
              if (!(h[par] < pes)) break;
              h[i] = h[par];
              i = par;
            }
            h[i] = pes;
            if (hn == k) {
              F = tol_upper(h[0], tol);
              T = sq_reject_bound(F);
            }
          } else if (pes < h[0]) {
            uint32_t i = 0;  // replace the largest, sift down
            for (;;) {
              const uint32_t l = 2 * i + 1, r = l + 1;
              uint32_t m = i;
              double vm = pes;
              if (l < k && vm < h[l]) m = l, vm = h[l];
              if (r < k && vm < h[r]) m = r;
              if (m == i) break;
              h[i] = h[m];
              i = m;
            }
            h[i] = pes;
            const double lim = tol_upper(h[0], tol);
            if (lim < F) {
              F = lim;
              T = sq_reject_bound(F);
            }
          }
        }
        for (uint32_t c = 0; c < ni.ccount; ++c) stack[top++] = ni.cbegin + c;
      }
    }
    Fq[qi] = F;
  });
}

// Heap sort of each band query's (distance, id) pairs: ascending distance for
// kNN, descending enter distance for rays; ties by id either way.
template <bool DESC>
__device__ __forceinline__ bool pair_less(double da, uint32_t ia, double db, uint32_t ib) {
  return (DESC ? da > db : da < db) || (da == db && ia < ib);
}

template <bool DESC, typename Off = uint32_t>
__global__ void __launch_bounds__(kQB) k_sort_band(const Off *__restrict__ offsets, uint64_t nband,
                                                    double *__restrict__ d, uint32_t *__restrict__ id) {
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nband) return;
  double *a = d + offsets[ti];
  uint32_t *b = id + offsets[ti];
  const uint32_t n = (uint32_t)(offsets[ti + 1] - offsets[ti]);
  auto sift = [&](uint32_t i, uint32_t m) {
    for (;;) {
      uint32_t l = 2 * i + 1, r = l + 1, x = i;
      if (l < m && pair_less<DESC>(a[x], b[x], a[l], b[l])) x = l;
      if (r < m && pair_less<DESC>(a[x], b[x], a[r], b[r])) x = r;
      if (x == i) return;
      const double td = a[i];
      a[i] = a[x];
      a[x] = td;
      const uint32_t tid = b[i];
      b[i] = b[x];
      b[x] = tid;
      i = x;
    }
  };
  for (uint32_t i = n / 2; i-- > 0;) sift(i, n);
  for (uint32_t m = n; m > 1; --m) {
    const double td = a[0];
    a[0] = a[m - 1];
    a[m - 1] = td;
    const uint32_t tid = b[0];
    b[0] = b[m - 1];
    b[m - 1] = tid;
    sift(0, m - 1);
  }
}

__global__ void k_band_counts(const uint32_t *__restrict__ band, uint32_t nband, const uint32_t *__restrict__ bcnt,
                              uint32_t *__restrict__ cnt) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < nband) cnt[band[i]] = bcnt[i];
}

__global__ void k_band_copy(const uint32_t *__restrict__ band, uint32_t nband, const uint32_t *__restrict__ boff,
                            const uint32_t *__restrict__ bid, const uint64_t *__restrict__ off,
                            uint32_t *__restrict__ dst) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nband) return;
  const uint64_t o = off[band[i]];
  for (uint32_t t = boff[i]; t < boff[i + 1]; ++t) dst[o + (t - boff[i])] = bid[t];
}

// Measurement helper: the nodes a kNN query must visit (wall distance below the
// query's final radius F) and the entities they hold, summed over the launch.
template <int D>
__global__ void __launch_bounds__(kQB) k_knn_footprint(const TreeView *__restrict__ views,
                                                        const uint32_t *__restrict__ qview,
                                                        const double *__restrict__ qpts,
                                                        const uint32_t *__restrict__ qset,
                                                        const uint32_t *__restrict__ order, uint64_t nq,
                                                        const double *__restrict__ Fq,
                                                        unsigned long long *__restrict__ acc) {
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  unsigned long long nodes = 0, ents = 0;
  if (ti < nq) {
    const uint64_t qi = order ? order[ti] : ti;
    with_tree(views, qview, qi, [&](const TreeView &tv) {
      const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
      if (sp.node_cnt == 0) return;
      const double F = Fq[qi];
      double q[D];
#pragma unroll
      for (int d = 0; d < D; ++d) q[d] = qpts[qi * D + d];
      uint32_t stack[StackCap<D>::value];
      int top = 0;
      stack[top++] = (uint32_t)sp.node_off;
      while (top > 0) {
        const uint32_t node = stack[--top];
        bool in;
        const double w2 = wall_distance2<D>(q, &tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &in);
        if (sqrt(w2) >= F) continue;
        const NodeInfo ni = gload(tv.info + node);
        ++nodes;
        ents += ni.ecount;
        for (uint32_t c = 0; c < ni.ccount; ++c) stack[top++] = ni.cbegin + c;
      }
    });
  }
  for (int o = 32; o > 0; o >>= 1) {
    nodes += __shfl_down(nodes, o);
    ents += __shfl_down(ents, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&acc[0], nodes);
    atomicAdd(&acc[1], ents);
  }
}

// ---------------------------------------------------------------------------
// range search
// ---------------------------------------------------------------------------
// internal_geometry_module.h:190-198 DoesRangeContainBox
__device__ __forceinline__ bool range_contains(double rmin, double rmax, double bmin, double bmax, double tol) {
  if (rmin - tol > bmin || bmin > rmax + tol) return false;
  if (rmin - tol > bmax || bmax > rmax + tol) return false;
  return true;
}

// adapters/general.h:262-289 GetBoxRelation(e1, e2, tol) == Overlapped
template <int D>
__device__ __forceinline__ bool boxes_overlap_strict(const double *mn1, const double *mx1, const double *mn2,
                                                     const double *mx2, double tol) {
  bool adjacent = false;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double a0 = mn1[d], a1 = mx1[d], b0 = mn2[d], b1 = mx2[d];
    if (a0 + tol < b1 && b0 + tol < a1) continue;
    if (fabs(a0 - b1) <= tol || fabs(a1 - b0) <= tol) {
      adjacent = true;
      continue;
    }
    if (b1 + tol < a0 || a1 + tol < b0) return false;
  }
  return !adjacent;
}

// Level stack in LDS for the range walk: per level with pending children, the
// next child and the end of the node's children.  A path holds at most one entry
// per tree depth, so the short stack never overflows.
constexpr int kRangeStack = kMaxDepth + 1;  // 20 x 5 B x 64 lanes = 6.25 KiB per block

struct LdsLevelIter {
  uint32_t *a;  // next pending child (children are taken last to first)
  uint8_t *b;   // children left at this level (<= 2^D)
  int stride = kQB;  // lanes of the block sharing the arrays
  __device__ __forceinline__ uint32_t &nx(int i) { return a[i * stride]; }
  __device__ __forceinline__ uint8_t &left(int i) { return b[i * stride]; }
};

// A contiguous run of ids a range result takes whole (a contained subtree, or
// 0..n-1 for the root shortcut when src is null).  Runs longer than kInlineCopy are
// copied afterwards by k_range_copy, one wavefront per run, instead of one lane.
// ids a range query's count walk also writes into its inline row; queries with no
// more results skip the fill walk
constexpr uint32_t kRangeInline = 8;

struct RangeCopy {
  uint64_t dst;
  const uint32_t *src;
  uint32_t len;
  uint32_t pad;
};
constexpr uint32_t kInlineCopy = 16;

// What a wavefront count walk keeps for the fill (collect mode): one record per node
// that adds ids -- a boundary node with the count of its entities that pass, a
// contained subtree, or the whole set -- at its offset in the query's list.  Listed
// query b (the b-th capped query) owns `cap` records, handed out by an LDS counter
// (no global atomics); k_range_kept then writes each record with a wavefront (the
// entity tests of a boundary node run again, the same tests, so the same ids).  The
// number of records per query, `cap`, is a batch's record budget over its capped
// queries, between kKeepMin and kKeepMax (or the context's range_records); records
// beyond it go to an overflow area shared by the batch (xcap = 1024 queries' slots, at
// most kKeepOverflow records, handed out by a global counter: only the few largest
// queries reach it).  A query whose records do not all fit is filled by the fill walk instead
// (ok[query] stays 0).
struct RangeRec {
  uint32_t node;
  uint32_t local;  // offset in the query's list
  uint32_t len;
  uint32_t kind;   // kRecNode, kRecSubtree, kRecSet
};
constexpr uint32_t kRecNode = 0, kRecSubtree = 1, kRecSet = 2;
struct RangeKeep {
  RangeRec *recs;  // cap per listed query
  uint32_t *nrec;  // per listed query: records in its own slots
  uint32_t *ok;    // per query (by query index): kept whole
  RangeRec *xrecs; // overflow area: xcap records
  uint32_t *xq;    // the listed query of each overflow record
  uint32_t *nx;    // overflow records handed out (zeroed before the walk)
  uint32_t cap, xcap;
  uint32_t nlist;  // listed queries with slots (those after them are not kept)
};
constexpr uint32_t kKeepMin = 256, kKeepMax = 8192, kKeepOverflow = 1u << 20;
constexpr uint64_t kKeepBudget = 16ull << 20;  // records per batch (256 MiB)

template <bool FILL>
__device__ __forceinline__ void range_take(uint32_t *__restrict__ out, uint64_t dst, const uint32_t *src, uint32_t len,
                                           uint32_t *__restrict__ ncopy, RangeCopy *__restrict__ copies) {
  if (len > kInlineCopy) {
    const uint32_t i = atomicAdd(ncopy, 1u);
    if (FILL) copies[i] = RangeCopy{dst, src, len, 0u};
  } else if (FILL) {
    for (uint32_t t = 0; t < len; ++t) out[dst + t] = src ? src[t] : t;
  }
}

// Entity test of RangeSearch (ot_query.h:543-580): points DoesBoxContainPoint
// (general.h:225-243); boxes Inside = AreBoxesOverlapped(range, e, true) (:296-308),
// Overlap = GetBoxRelation == Overlapped.
template <int D, bool BOX>
__device__ __forceinline__ bool range_entity(const TreeView &tv, uint32_t j, const double (&rmin)[D],
                                             const double (&rmax)[D], int overlap_mode, double tol) {
  bool ok = true;
  if (!BOX) {
    if (tol != 0.0) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double p = tv.geom[(uint64_t)d * tv.n + j];
        ok &= (rmin[d] - tol < p && p < rmax[d] + tol);
      }
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double p = tv.geom[(uint64_t)d * tv.n + j];
        ok &= (rmin[d] <= p && p <= rmax[d]);
      }
    }
  } else {
    double emn[D], emx[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      emn[d] = tv.geom[(uint64_t)d * tv.n + j];
      emx[d] = tv.geom[(uint64_t)(D + d) * tv.n + j];
    }
    if (overlap_mode) {
      ok = boxes_overlap_strict<D>(rmin, rmax, emn, emx, tol);
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (rmin[d] > emn[d] || emn[d] > rmax[d]) ok = false;
        if (rmin[d] > emx[d] || emx[d] > rmax[d]) ok = false;
      }
    }
  }
  return ok;
}

// Node test of the range walk: AreBoxesOverlappingByMinPoint(range, node)
// (internal_geometry_module.h:148-161) and DoesRangeContainBoxAD (:201-252).
template <int D>
__device__ __forceinline__ void range_node_test(const TreeView &tv, uint32_t node, const double (&rmin)[D],
                                                const double (&rmax)[D], const double (&rsz)[D], double tol,
                                                bool &overlap, bool &contained) {
  const double *nm = &tv.nmin[(uint64_t)node * D];
  const double *ns = &tv.nsize[(uint64_t)node * D];
  overlap = true;
  contained = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double maxL = rmin[d] + rsz[d];
    const double maxR = nm[d] + ns[d];
    if (maxL <= nm[d] - tol || maxR <= rmin[d] - tol) overlap = false;
    contained &= range_contains(rmin[d], rmax[d], nm[d], nm[d] + ns[d], tol);
  }
}

// count of a query whose one-lane count walk stopped at the visit cap
constexpr uint32_t kRangeCapped = 0xFFFFFFFFu;

template <int D, bool BOX, bool FILL>
__device__ __forceinline__ void range_query(const TreeView &tv, uint64_t qi, const double *__restrict__ ranges,
                                            const uint32_t *__restrict__ qset, int overlap_mode, double tol,
                                            LdsLevelIter &stack, uint32_t *__restrict__ counts,
                                            const uint64_t *__restrict__ offsets, uint32_t *__restrict__ out,
                                            uint32_t *__restrict__ ncopy, RangeCopy *__restrict__ copies,
                                            uint32_t *__restrict__ inl, uint32_t vcap) {
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  double rmin[D], rmax[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    rmin[d] = ranges[qi * 2 * D + d];
    rmax[d] = ranges[qi * 2 * D + D + d];
  }
  uint64_t o = FILL ? offsets[qi] : 0;
  uint32_t cnt = 0, visits = 0;
  bool capped = false;
  if (sp.node_cnt != 0) {
    const uint32_t root = (uint32_t)sp.node_off;
    bool root_in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double mn = tv.nmin[(uint64_t)root * D + d];
      const double bmax = mn + tv.nsize[(uint64_t)root * D + d];
      root_in &= range_contains(rmin[d], rmax[d], mn, bmax, tol);
    }
    double vol = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) vol *= rmax[d] - rmin[d];
    if (root_in) {
      cnt = (uint32_t)sp.ent_cnt;  // iota over the entity span (:488-512)
      range_take<FILL>(out, o, nullptr, cnt, ncopy, copies);
      if (!FILL && inl && cnt <= kRangeInline)
        for (uint32_t t = 0; t < cnt; ++t) inl[t] = t;
    } else if (vol > 0.0) {
      double rsz[D];
#pragma unroll
      for (int d = 0; d < D; ++d) rsz[d] = rmax[d] - rmin[d];
      // walk: one stack entry per level with pending children (next, end), so the
      // depth never exceeds the tree's
      int top = 0;
      uint32_t node = root;
      for (;;) {
        if (!FILL && vcap && ++visits > vcap) {  // a long walk: left to a wavefront
          capped = true;
          break;
        }
        const NodeInfo ni = gload(tv.info + node);  // loaded with the box that decides the visit
        bool overlap, contained;
        range_node_test<D>(tv, node, rmin, rmax, rsz, tol, overlap, contained);
        if (overlap && contained) {
          // whole subtree: its entities are the contiguous storage range
          const uint32_t e0 = ni.ebegin, e1 = tv.rend[node];
          range_take<FILL>(out, o + cnt, tv.ids + e0, e1 - e0, ncopy, copies);
          if (!FILL && inl && cnt + (e1 - e0) <= kRangeInline)
            for (uint32_t t = 0; t < e1 - e0; ++t) inl[cnt + t] = tv.ids[e0 + t];
          cnt += e1 - e0;
        } else if (overlap) {
          for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
            if (range_entity<D, BOX>(tv, j, rmin, rmax, overlap_mode, tol)) {
              if (FILL) out[o + cnt] = tv.ids[j];
              else if (inl && cnt < kRangeInline) inl[cnt] = tv.ids[j];
              ++cnt;
            }
          }
          if (ni.ccount) {  // last child first, the reference's depth-first output order
            stack.nx(top) = ni.cbegin + ni.ccount - 1;
            stack.left(top) = (uint8_t)ni.ccount;
            ++top;
          }
        }
        if (top == 0) break;
        node = stack.nx(top - 1)--;
        if (--stack.left(top - 1) == 0) --top;
      }
    }
  }
  if (!FILL) counts[qi] = capped ? kRangeCapped : cnt;
}

template <int D, bool BOX, bool FILL>
__global__ void __launch_bounds__(kQB) k_range(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                                const double *__restrict__ ranges, const uint32_t *__restrict__ qset,
                                                const uint32_t *__restrict__ order, uint64_t nq, int overlap_mode,
                                                double tol, uint32_t *__restrict__ counts,
                                                const uint64_t *__restrict__ offsets, uint32_t *__restrict__ out,
                                                uint32_t *__restrict__ ncopy, RangeCopy *__restrict__ copies,
                                                uint32_t *__restrict__ inl, uint32_t vcap) {
  __shared__ uint32_t s_a[kRangeStack * kQB];
  __shared__ uint8_t s_b[kRangeStack * kQB];
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nq) return;
  const uint64_t qi = order ? order[ti] : ti;
  LdsLevelIter S{&s_a[threadIdx.x], &s_b[threadIdx.x]};
  uint32_t *row = inl ? inl + ti * kRangeInline : nullptr;
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    range_query<D, BOX, FILL>(tv, qi, ranges, qset, overlap_mode, tol, S, counts, offsets, out, ncopy, copies, row,
                              vcap);
  });
}

// One wavefront per range query of a list (the queries the one-lane walk leaves:
// more than kRangeInline ids, or more node visits than its cap): the walk of
// range_query with the frontier expansion and per-lane subtree walks of
// k_planes_wave.  Every node reached has all its ancestors visited, so the id set is
// range_query's; ids are appended at the query's CSR offset in no particular order.
template <int D, bool BOX, bool FILL>
__global__ void __launch_bounds__(64) k_range_wave(const TreeView *__restrict__ views,
                                                    const uint32_t *__restrict__ qview,
                                                    const double *__restrict__ ranges,
                                                    const uint32_t *__restrict__ qset,
                                                    const uint32_t *__restrict__ list, int overlap_mode, double tol,
                                                    uint32_t *__restrict__ counts,
                                                    const uint64_t *__restrict__ offsets, uint32_t *__restrict__ out,
                                                    uint32_t *__restrict__ ncopy, RangeCopy *__restrict__ copies,
                                                    const uint32_t *__restrict__ nlist,
                                                    const uint32_t *__restrict__ abort, const RangeKeep keep) {
  static_assert(D <= 3, "frontier sized for up to 8 children per node");
  // one frontier, rewritten in place each round: fewer than 64 nodes of <= 8 children
  __shared__ uint32_t s_front[64 * 8];
  __shared__ uint32_t s_nf[2], s_cnt, s_next, s_nrec, s_bad;
  __shared__ uint32_t s_a[kRangeStack * 64];
  __shared__ uint8_t s_b[kRangeStack * 64];
  if (abort && *abort) return;
  const uint32_t lane = threadIdx.x;
  // list entries: one per block, or (a device-side count) strided over a bounded grid
  const uint32_t nl = nlist ? *nlist : gridDim.x;
  for (uint32_t b = blockIdx.x; b < nl; b += gridDim.x) {
  const uint64_t qi = list[b];
  const TreeView &tv = views[qview ? qview[qi] : 0u];
  const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
  double rmin[D], rmax[D], rsz[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    rmin[d] = ranges[qi * 2 * D + d];
    rmax[d] = ranges[qi * 2 * D + D + d];
    rsz[d] = rmax[d] - rmin[d];
  }
  const uint64_t o = FILL ? offsets[qi] : 0;
  // 0 nothing, 1 the whole set (root contained: iota, :488-512), 2 walk
  int start = 0;
  if (sp.node_cnt != 0) {
    const uint32_t root = (uint32_t)sp.node_off;
    bool root_in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double mn = tv.nmin[(uint64_t)root * D + d];
      root_in &= range_contains(rmin[d], rmax[d], mn, mn + tv.nsize[(uint64_t)root * D + d], tol);
    }
    double vol = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) vol *= rmax[d] - rmin[d];
    start = root_in ? 1 : (vol > 0.0 ? 2 : 0);
  }
  // collect mode (count walk with `keep`): slots of listed query b
  const bool kept = !FILL && keep.recs && b < keep.nlist;
  RangeRec *const krecs = kept ? keep.recs + (uint64_t)b * keep.cap : nullptr;
  auto keep_rec = [&](uint32_t node, uint32_t at, uint32_t len, uint32_t kind) {
    const uint32_t r = atomicAdd(&s_nrec, 1u);
    if (r < keep.cap) {
      krecs[r] = RangeRec{node, at, len, kind};
    } else {
      const uint32_t x = atomicAdd(keep.nx, 1u);
      if (x < keep.xcap) {
        keep.xrecs[x] = RangeRec{node, at, len, kind};
        keep.xq[x] = b;
      } else {
        s_bad = 1;
      }
    }
  };
  auto emit_run = [&](const uint32_t *src, uint32_t len, uint32_t node, uint32_t kind) {
    const uint32_t at = atomicAdd(&s_cnt, len);
    range_take<FILL>(out, o + at, src, len, ncopy, copies);  // counted also when kept: the bound of a fill walk
    if (kept) keep_rec(node, at, len, kind);
  };
  // node visit: whole subtree, or entities + children (returned through ni)
  auto visit = [&](uint32_t node, NodeInfo &ni) -> bool {
    bool overlap, contained;
    range_node_test<D>(tv, node, rmin, rmax, rsz, tol, overlap, contained);
    if (!overlap) return false;
    ni = tv.info[node];
    if (contained) {
      emit_run(tv.ids + ni.ebegin, tv.rend[node] - ni.ebegin, node, kRecSubtree);
      return false;
    }
    if (FILL) {
      for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j)
        if (range_entity<D, BOX>(tv, j, rmin, rmax, overlap_mode, tol)) out[o + atomicAdd(&s_cnt, 1u)] = tv.ids[j];
    } else {
      uint32_t c = 0;
      for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j)
        c += range_entity<D, BOX>(tv, j, rmin, rmax, overlap_mode, tol) ? 1u : 0u;
      if (c) {
        const uint32_t at = atomicAdd(&s_cnt, c);
        if (kept) keep_rec(node, at, c, kRecNode);
      }
    }
    return ni.ccount != 0;
  };
  if (lane == 0) {
    s_cnt = 0;
    s_nrec = 0;
    s_bad = 0;
    s_next = 0;
    s_nf[0] = 0;
    s_nf[1] = 0;
    if (start == 1) emit_run(nullptr, (uint32_t)sp.ent_cnt, 0u, kRecSet);
    if (start == 2) {
      s_front[0] = (uint32_t)sp.node_off;
      s_nf[0] = 1;
    }
  }
  __syncthreads();
  int cur = 0;
  while (s_nf[cur] != 0 && s_nf[cur] < 64) {
    const uint32_t nf = s_nf[cur];
    const uint32_t mine = lane < nf ? s_front[lane] : 0u;  // read before the buffer is rewritten
    if (lane == 0) s_nf[cur ^ 1] = 0;
    __syncthreads();
    if (lane < nf) {
      NodeInfo ni;
      if (visit(mine, ni)) {
        const uint32_t at = atomicAdd(&s_nf[cur ^ 1], ni.ccount);
        for (uint32_t c = 0; c < ni.ccount; ++c) s_front[at + c] = ni.cbegin + c;
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  const uint32_t nf = s_nf[cur];
  LdsLevelIter stack{&s_a[lane], &s_b[lane], 64};
  for (;;) {
    const uint32_t item = atomicAdd(&s_next, 1u);
    if (item >= nf) break;
    int top = 0;
    uint32_t node = s_front[item];
    for (;;) {
      NodeInfo ni;
      if (visit(node, ni)) {
        stack.nx(top) = ni.cbegin + ni.ccount - 1;
        stack.left(top) = (uint8_t)ni.ccount;
        ++top;
      }
      if (top == 0) break;
      node = stack.nx(top - 1)--;
      if (--stack.left(top - 1) == 0) --top;
    }
  }
  __syncthreads();
  if (!FILL && lane == 0) {
    counts[qi] = s_cnt;
    if (!FILL && keep.recs) {
      if (kept) keep.nrec[b] = min(s_nrec, keep.cap);
      keep.ok[qi] = (kept && !s_bad) ? 1u : 0u;
    }
  }
  }
}

// One kept record into the CSR, by a wavefront.  A boundary node's entities are tested
// again 64 at a time; the ones that pass are placed by their rank in the ballot, so the
// record's ids fill exactly its `len` slots.
template <int D, bool BOX>
__device__ __forceinline__ void range_kept_rec(const TreeView &tv, const RangeRec rec, const double (&rmin)[D],
                                               const double (&rmax)[D], int overlap_mode, double tol,
                                               uint32_t *__restrict__ dst, uint32_t lane) {
  if (rec.kind == kRecSet) {
    for (uint32_t t = lane; t < rec.len; t += 64) dst[t] = t;
  } else if (rec.kind == kRecSubtree) {
    const uint32_t *src = tv.ids + tv.info[rec.node].ebegin;
    for (uint32_t t = lane; t < rec.len; t += 64) dst[t] = src[t];
  } else {
    const NodeInfo ni = tv.info[rec.node];
    uint32_t at = 0;
    for (uint32_t base = 0; base < ni.ecount; base += 64) {
      const uint32_t j = ni.ebegin + base + lane;
      const bool pass = base + lane < ni.ecount && range_entity<D, BOX>(tv, j, rmin, rmax, overlap_mode, tol);
      const uint64_t m = __ballot(pass);
      if (pass) dst[at + __popcll(m & ((1ull << lane) - 1))] = tv.ids[j];
      at += __popcll(m);
    }
  }
}

// Kept records into the CSR: a block per listed query (its own slots) and, strided over
// the rest of the grid, the overflow records; a wavefront per record.
template <int D, bool BOX>
__global__ void __launch_bounds__(256) k_range_kept(const TreeView *__restrict__ views,
                                                     const uint32_t *__restrict__ qview,
                                                     const double *__restrict__ ranges,
                                                     const uint32_t *__restrict__ list, uint32_t nlist,
                                                     int overlap_mode, double tol, const RangeKeep keep,
                                                     const uint64_t *__restrict__ off, uint32_t *__restrict__ out,
                                                     const uint32_t *__restrict__ ncount,
                                                     const uint32_t *__restrict__ abort) {
  // ncount (when given): the device-side length of the list; abort: results over a
  // plan's capacity, nothing is written
  if (abort && *abort) return;
  const uint32_t lane = threadIdx.x & 63;
  auto load = [&](uint32_t b, uint64_t &qi, double (&rmin)[D], double (&rmax)[D]) {
    qi = list[b];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      rmin[d] = ranges[qi * 2 * D + d];
      rmax[d] = ranges[qi * 2 * D + D + d];
    }
  };
  double rmin[D], rmax[D];
  uint64_t qi;
  if (blockIdx.x < nlist) {
    const uint32_t b = blockIdx.x;
    if (ncount && b >= *ncount) return;
    load(b, qi, rmin, rmax);
    if (!keep.ok[qi]) return;
    const TreeView &tv = views[qview ? qview[qi] : 0u];
    const uint64_t o = off[qi];
    const RangeRec *recs = keep.recs + (uint64_t)b * keep.cap;
    const uint32_t nrec = keep.nrec[b];
    for (uint32_t r = threadIdx.x / 64; r < nrec; r += 4)
      range_kept_rec<D, BOX>(tv, recs[r], rmin, rmax, overlap_mode, tol, out + o + recs[r].local, lane);
    return;
  }
  const uint32_t nx = min(*keep.nx, keep.xcap);
  const uint32_t waves = (gridDim.x - nlist) * 4;
  for (uint32_t x = (blockIdx.x - nlist) * 4 + threadIdx.x / 64; x < nx; x += waves) {
    load(keep.xq[x], qi, rmin, rmax);
    if (!keep.ok[qi]) continue;
    const TreeView &tv = views[qview ? qview[qi] : 0u];
    const RangeRec rec = keep.xrecs[x];
    range_kept_rec<D, BOX>(tv, rec, rmin, rmax, overlap_mode, tol, out + off[qi] + rec.local, lane);
  }
}

// Queries whose count walk found at most kRangeInline ids have them all in their
// row of the inline buffer (walk order): copied into the CSR here, one thread per
// (query, slot).  The others are flagged for the fill walk.
// (ids ascending: every range result is returned sorted by id)
__global__ void k_range_inline(const uint32_t *__restrict__ order, uint64_t nq, const uint32_t *__restrict__ cnt,
                               const uint32_t *__restrict__ capped, const uint64_t *__restrict__ off,
                               const uint32_t *__restrict__ inl, uint32_t *__restrict__ out,
                               uint32_t *__restrict__ heavy, const uint32_t *__restrict__ abort,
                               uint32_t *__restrict__ walk, const uint32_t *__restrict__ ok) {
  const uint64_t ti = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (ti >= nq) return;
  if (abort && *abort) {  // results over the plan's capacity: nothing is written
    heavy[ti] = 0;
    return;
  }
  const uint64_t qi = order ? order[ti] : ti;
  const uint32_t c = cnt[qi];
  const bool h = capped[ti] || c > kRangeInline;  // a capped walk left its row incomplete
  heavy[ti] = h;
  // walk (when given): the heavy queries the fill walk still has to fill -- not the
  // capped ones whose count walk kept their whole lists (ok; null: none kept)
  if (walk) walk[ti] = h && !(capped[ti] && ok && ok[qi]);
  if (h || c == 0) return;
  uint32_t v[kRangeInline];
#pragma unroll
  for (uint32_t t = 0; t < kRangeInline; ++t) v[t] = t < c ? inl[ti * kRangeInline + t] : 0xFFFFFFFFu;
#pragma unroll
  for (uint32_t a = 1; a < kRangeInline; ++a)
#pragma unroll
    for (uint32_t b = a; b > 0; --b)
      if (v[b] < v[b - 1]) {
        const uint32_t tmp = v[b];
        v[b] = v[b - 1];
        v[b - 1] = tmp;
      }
  const uint64_t o = off[qi];
#pragma unroll
  for (uint32_t t = 0; t < kRangeInline; ++t)
    if (t < c) out[o + t] = v[t];
}

// Sorting the larger results: key (rank of the query among them, id) per id.
// (nh: a device-side count of listed queries, strided over a bounded grid; abort: skip)
__global__ void k_range_sort_keys(const uint32_t *__restrict__ hlist, uint32_t nheavy, const uint64_t *__restrict__ off,
                                  const uint32_t *__restrict__ hpos, const uint32_t *__restrict__ ids,
                                  uint64_t *__restrict__ keys, const uint32_t *__restrict__ nh,
                                  const uint32_t *__restrict__ abort, uint32_t idbits) {
  if (abort && *abort) return;
  const uint32_t n = nh ? *nh : nheavy;
  for (uint32_t h = blockIdx.x; h < n; h += gridDim.x) {
    const uint32_t qi = hlist[h];
    const uint64_t o = off[qi], c = off[qi + 1] - o;
    for (uint64_t t = threadIdx.x; t < c; t += blockDim.x) keys[hpos[h] + t] = ((uint64_t)h << idbits) | ids[o + t];
  }
}

__global__ void k_range_sort_back(const uint32_t *__restrict__ hlist, uint32_t nheavy, const uint64_t *__restrict__ off,
                                  const uint32_t *__restrict__ hpos, const uint64_t *__restrict__ keys,
                                  uint32_t *__restrict__ ids, const uint32_t *__restrict__ nh,
                                  const uint32_t *__restrict__ abort, uint32_t idbits) {
  if (abort && *abort) return;
  const uint32_t n = nh ? *nh : nheavy;
  const uint64_t mask = idbits >= 32 ? 0xFFFFFFFFull : ((1ull << idbits) - 1);
  for (uint32_t h = blockIdx.x; h < n; h += gridDim.x) {
    const uint32_t qi = hlist[h];
    const uint64_t o = off[qi], c = off[qi + 1] - o;
    for (uint64_t t = threadIdx.x; t < c; t += blockDim.x) ids[o + t] = (uint32_t)(keys[hpos[h] + t] & mask);
  }
}

__global__ void k_iota_u32(uint32_t *__restrict__ a, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = i;
}

__global__ void k_range_heavy_counts(const uint32_t *__restrict__ hlist, uint32_t nheavy,
                                     const uint64_t *__restrict__ off, uint32_t *__restrict__ hc,
                                     const uint32_t *__restrict__ nh) {
  const uint64_t h = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (h < (nh ? *nh : nheavy)) hc[h] = (uint32_t)(off[hlist[h] + 1] - off[hlist[h]]);
}

__global__ void k_range_capped(const uint32_t *__restrict__ order, uint64_t nq, const uint32_t *__restrict__ cnt,
                               uint32_t *__restrict__ capped) {
  const uint64_t ti = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (ti < nq) capped[ti] = cnt[order ? order[ti] : ti] == kRangeCapped;
}

__global__ void k_range_heavy_list(const uint32_t *__restrict__ heavy, const uint32_t *__restrict__ pos, uint64_t nq,
                                   const uint32_t *__restrict__ order, uint32_t *__restrict__ horder) {
  const uint64_t ti = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (ti < nq && heavy[ti]) horder[pos[ti]] = order ? order[ti] : (uint32_t)ti;
}

// One wavefront per deferred run of a range result.
// (ncopy: a device-side run count, strided over a bounded grid; abort: skip)
__global__ void __launch_bounds__(256) k_range_copy(const RangeCopy *__restrict__ copies, uint32_t n,
                                                     uint32_t *__restrict__ out, const uint32_t *__restrict__ ncopy,
                                                     const uint32_t *__restrict__ abort) {
  if (abort && *abort) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t total = ncopy ? *ncopy : n;
  for (uint32_t w = blockIdx.x * 4 + threadIdx.x / 64; w < total; w += gridDim.x * 4) {
    const RangeCopy c = copies[w];
    for (uint32_t t = lane; t < c.len; t += 64) out[c.dst + t] = c.src ? c.src[t] : t;
  }
}

// ---------------------------------------------------------------------------
// planes: PlaneSearch, PlanePositiveSegmentation, FrustumCulling
// ---------------------------------------------------------------------------
// internal_geometry_module.h:254-280 GetBoxPlaneRelationAD(minPoint, size, d, n, tol):
// -1 negative side, 0 hit, 1 positive side.  pl = (origo distance, normal[D]).
template <int D>
__device__ __forceinline__ int box_plane_relation(const double (&mn)[D], const double (&sz)[D], const double *pl,
                                                  double tol) {
  double half[D], center[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    half[d] = sz[d] * 0.5;
    center[d] = mn[d] + half[d];
  }
  double radius = tol;
#pragma unroll
  for (int d = 0; d < D; ++d) radius += half[d] * fabs(pl[1 + d]);
  double dot = 0.0;  // DotAD(planeNormal, center) :180-188
#pragma unroll
  for (int d = 0; d < D; ++d) dot += pl[1 + d] * center[d];
  const double cp = dot - pl[0];
  if (cp + radius < 0.0) return -1;
  if (cp - radius > 0.0) return 1;
  return 0;
}

// ot_query.h:585-602 GetEntityPlaneRelation: points GetPointPlaneRelation
// (general.h:322-335), boxes GetBoxPlaneRelationAD(min, max - min)
template <int D, bool BOX>
__device__ __forceinline__ int entity_plane_relation(const TreeView &tv, uint32_t j, const double *pl, double tol) {
  if (BOX) {
    double mn[D], sz[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      mn[d] = tv.geom[(uint64_t)d * tv.n + j];
      sz[d] = tv.geom[(uint64_t)(D + d) * tv.n + j] - mn[d];
    }
    return box_plane_relation<D>(mn, sz, pl, tol);
  }
  double proj = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) proj += pl[1 + d] * tv.geom[(uint64_t)d * tv.n + j];
  if (proj < pl[0] - tol) return -1;
  if (proj > pl[0] + tol) return 1;
  return 0;
}

// MODE 0 PlaneSearch (ot_query.h:612-636): nodes and entities the plane hits.
// MODE 1 PlanePositiveSegmentation (:691-741): a node wholly on the positive side
// gives its whole subtree (a contiguous storage range), a hit node its entities not
// on the negative side.  MODE 2 FrustumCulling (:756-822): a node is visited unless a
// plane met before any hitting plane has it on the negative side; entities on no
// plane's negative side.  Depth-first, last child first (the reference's order).
template <int D, bool BOX, int MODE>
__device__ __forceinline__ bool planes_node(const TreeView &tv, uint32_t node, const double *pq, uint32_t nplanes,
                                            double tol, NodeInfo &ni, bool &whole);
template <int D, bool BOX, int MODE>
__device__ __forceinline__ bool planes_entity(const TreeView &tv, uint32_t j, const double *pq, uint32_t nplanes,
                                              double tol);

template <int D, bool BOX, int MODE, bool FILL>
__device__ __forceinline__ void planes_query(const TreeView &tv, uint64_t qi, const double *__restrict__ planes,
                                             uint32_t nplanes, const uint32_t *__restrict__ qset, double tol,
                                             LdsLevelIter &stack, uint32_t *__restrict__ counts,
                                             const uint64_t *__restrict__ offsets, uint32_t *__restrict__ out,
                                             uint32_t *__restrict__ ncopy, RangeCopy *__restrict__ copies) {
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  const double *pq = planes + qi * (uint64_t)nplanes * (D + 1);
  const uint64_t o = FILL ? offsets[qi] : 0;
  uint32_t cnt = 0;
  if (sp.node_cnt != 0 && nplanes != 0) {
    int top = 0;
    uint32_t node = (uint32_t)sp.node_off;
    for (;;) {
      NodeInfo ni;
      bool whole;
      if (planes_node<D, BOX, MODE>(tv, node, pq, nplanes, tol, ni, whole)) {
        if (whole) {
          const uint32_t e0 = ni.ebegin, e1 = tv.rend[node];
          range_take<FILL>(out, o + cnt, tv.ids + e0, e1 - e0, ncopy, copies);
          cnt += e1 - e0;
        } else {
          for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j)
            if (planes_entity<D, BOX, MODE>(tv, j, pq, nplanes, tol)) {
              if (FILL) out[o + cnt] = tv.ids[j];
              ++cnt;
            }
          if (ni.ccount) {
            stack.nx(top) = ni.cbegin + ni.ccount - 1;
            stack.left(top) = (uint8_t)ni.ccount;
            ++top;
          }
        }
      }
      if (top == 0) break;
      node = stack.nx(top - 1)--;
      if (--stack.left(top - 1) == 0) --top;
    }
  }
  if (!FILL) counts[qi] = cnt;
}

template <int D, bool BOX, int MODE, bool FILL>
__global__ void __launch_bounds__(kQB) k_planes(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                                 const double *__restrict__ planes, uint32_t nplanes,
                                                 const uint32_t *__restrict__ qset, uint64_t nq, double tol,
                                                 uint32_t *__restrict__ counts, const uint64_t *__restrict__ offsets,
                                                 uint32_t *__restrict__ out, uint32_t *__restrict__ ncopy,
                                                 RangeCopy *__restrict__ copies) {
  __shared__ uint32_t s_a[kRangeStack * kQB];
  __shared__ uint8_t s_b[kRangeStack * kQB];
  const uint64_t qi = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (qi >= nq) return;
  LdsLevelIter S{&s_a[threadIdx.x], &s_b[threadIdx.x]};
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    planes_query<D, BOX, MODE, FILL>(tv, qi, planes, nplanes, qset, tol, S, counts, offsets, out, ncopy, copies);
  });
}

// Node test and entity pass of one visited node of a plane query, shared by the
// one-lane walk and the wavefront walk.  Returns whether the walk descends into the
// node's children; ids go through emit(count or id source) -- appended ids or whole
// subtree runs (MODE 1, a node wholly on the positive side).
template <int D, bool BOX, int MODE>
__device__ __forceinline__ bool planes_node(const TreeView &tv, uint32_t node, const double *pq, uint32_t nplanes,
                                            double tol, NodeInfo &ni, bool &whole) {
  double mn[D], sz[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    mn[d] = tv.nmin[(uint64_t)node * D + d];
    sz[d] = tv.nsize[(uint64_t)node * D + d];
  }
  bool visit = true;
  whole = false;
  if (MODE == 2) {
    for (uint32_t k = 0; k < nplanes; ++k) {
      const int rel = box_plane_relation<D>(mn, sz, pq + k * (D + 1), tol);
      if (rel == 0) break;
      if (rel < 0) {
        visit = false;
        break;
      }
    }
  } else {
    const int rel = box_plane_relation<D>(mn, sz, pq, tol);
    if (MODE == 0) visit = rel == 0;
    else {
      visit = rel >= 0;
      whole = rel > 0;
    }
  }
  if (visit) ni = tv.info[node];
  return visit;
}

template <int D, bool BOX, int MODE>
__device__ __forceinline__ bool planes_entity(const TreeView &tv, uint32_t j, const double *pq, uint32_t nplanes,
                                              double tol) {
  if (MODE == 0) return entity_plane_relation<D, BOX>(tv, j, pq, tol) == 0;
  bool ok = true;
  for (uint32_t k = 0; k < nplanes && ok; ++k) ok = entity_plane_relation<D, BOX>(tv, j, pq + k * (D + 1), tol) >= 0;
  return ok;
}

// One wavefront per plane query: a plane or frustum meets a whole slab of the
// tree, so a query is thousands of node visits and a launch holds few queries.
// Phase 1: the wavefront expands the top of the tree breadth-first (one lane per
// frontier node, the node test and entity pass as in planes_query) until the
// frontier holds at least W nodes or nothing is left to expand; a round starts
// from fewer than W nodes of at most 2^D children, so the frontier holds at most
// (W - 1) * 2^D entries and fits s_front[W * 8].  Phase 2: each lane takes frontier nodes
// from an LDS counter and walks their subtrees with its own level iterators (the
// walk of planes_query, bounded by the tree depth).  Every node a walk reaches has
// all its ancestors visited, so the set of ids is planes_query's; ids are appended
// at the query's CSR offset through an LDS counter, in no particular order.
template <int D, bool BOX, int MODE, bool FILL, int W = 64>
__global__ void __launch_bounds__(W) k_planes_wave(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                                     const double *__restrict__ planes, uint32_t nplanes,
                                                     const uint32_t *__restrict__ qset, uint64_t nq, double tol,
                                                     uint32_t *__restrict__ counts, const uint64_t *__restrict__ offsets,
                                                     uint32_t *__restrict__ out, uint32_t *__restrict__ ncopy,
                                                     RangeCopy *__restrict__ copies) {
  static_assert(D <= 3, "frontier sized for up to 8 children per node");
  // one frontier, rewritten in place each round; it grows while narrower than the
  // block: fewer than W nodes of <= 8 children
  __shared__ uint32_t s_front[W * 8];
  __shared__ uint32_t s_nf[2], s_cnt, s_next;
  __shared__ uint32_t s_a[kRangeStack * W];
  __shared__ uint8_t s_b[kRangeStack * W];
  const uint64_t qi = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const TreeView &tv = views[qview ? qview[qi] : 0u];
  const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
  const double *pq = planes + qi * (uint64_t)nplanes * (D + 1);
  const uint64_t o = FILL ? offsets[qi] : 0;
  auto emit_run = [&](const uint32_t *src, uint32_t len) {
    const uint32_t at = atomicAdd(&s_cnt, len);
    range_take<FILL>(out, o + at, src, len, ncopy, copies);
  };
  auto entities = [&](const NodeInfo &ni) {
    for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j)
      if (planes_entity<D, BOX, MODE>(tv, j, pq, nplanes, tol)) {
        const uint32_t at = atomicAdd(&s_cnt, 1u);
        if (FILL) out[o + at] = tv.ids[j];
      }
  };
  if (lane == 0) {
    s_cnt = 0;
    s_next = 0;
    s_nf[0] = 0;
    s_nf[1] = 0;
    if (sp.node_cnt != 0 && nplanes != 0) {
      s_front[0] = (uint32_t)sp.node_off;
      s_nf[0] = 1;
    }
  }
  __syncthreads();
  // phase 1: breadth-first expansion while the frontier is narrower than the block
  int cur = 0;
  while (s_nf[cur] != 0 && s_nf[cur] < W) {
    const uint32_t nf = s_nf[cur];
    const uint32_t mine = lane < nf ? s_front[lane] : 0u;  // read before the buffer is rewritten
    if (lane == 0) s_nf[cur ^ 1] = 0;
    __syncthreads();
    if (lane < nf) {
      const uint32_t node = mine;
      NodeInfo ni;
      bool whole;
      if (planes_node<D, BOX, MODE>(tv, node, pq, nplanes, tol, ni, whole)) {
        if (whole) emit_run(tv.ids + ni.ebegin, tv.rend[node] - ni.ebegin);
        else {
          entities(ni);
          if (ni.ccount) {
            const uint32_t at = atomicAdd(&s_nf[cur ^ 1], ni.ccount);
            for (uint32_t c = 0; c < ni.ccount; ++c) s_front[at + c] = ni.cbegin + c;
          }
        }
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  // phase 2: the lanes walk the frontier's subtrees
  const uint32_t nf = s_nf[cur];
  LdsLevelIter stack{&s_a[lane], &s_b[lane], W};
  for (;;) {
    const uint32_t item = atomicAdd(&s_next, 1u);
    if (item >= nf) break;
    int top = 0;
    uint32_t node = s_front[item];
    for (;;) {
      NodeInfo ni;
      bool whole;
      if (planes_node<D, BOX, MODE>(tv, node, pq, nplanes, tol, ni, whole)) {
        if (whole) emit_run(tv.ids + ni.ebegin, tv.rend[node] - ni.ebegin);
        else {
          entities(ni);
          if (ni.ccount) {
            stack.nx(top) = ni.cbegin + ni.ccount - 1;
            stack.left(top) = (uint8_t)ni.ccount;
            ++top;
          }
        }
      }
      if (top == 0) break;
      node = stack.nx(top - 1)--;
      if (--stack.left(top - 1) == 0) --top;
    }
  }
  __syncthreads();
  if (!FILL && lane == 0) counts[qi] = s_cnt;
}

// ---------------------------------------------------------------------------
// pick search
// ---------------------------------------------------------------------------
// internal_geometry_module.h:372-396 DoesBoxContainPointAD(box, point, tol) and
// adapters/general.h:225-243 DoesBoxContainPoint (the same comparisons)
template <int D>
__device__ __forceinline__ bool box_contains_point(const double *mn, const double *mx, const double (&p)[D], double tol) {
  bool ok = true;
  if (tol != 0.0) {
#pragma unroll
    for (int d = 0; d < D; ++d) ok &= (mn[d] - tol < p[d] && p[d] < mx[d] + tol);
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d) ok &= (mn[d] <= p[d] && p[d] <= mx[d]);
  }
  return ok;
}

// internal_geometry_module.h:398-422 DoesBoxContainPointAD(minPoint, size, point, tol):
// |p - min - size/2| against size/2 (+ tol)
template <int D>
__device__ __forceinline__ bool node_contains_point(const double *mn, const double *sz, const double (&p)[D], double tol) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double half = sz[d] * 0.5;
    const double dist = fabs(p[d] - mn[d] - half);
    ok &= (tol != 0.0) ? !(dist >= half + tol) : !(dist > half);
  }
  return ok;
}

// PickSearch (ot_query.h:426-466): root box test, then depth-first (last child
// first, the reference's output order) skipping nodes whose box misses the point.
// Entities: boxes containing the point, or points within tol (Distance2 <= tol^2,
// general.h:218-221).  Count pass (FILL = false) and fill pass.
template <int D, bool BOX, bool FILL>
__device__ __forceinline__ void pick_query(const TreeView &tv, uint64_t qi, const double *__restrict__ picks,
                                           const uint32_t *__restrict__ qset, double tol, LdsLevelIter &stack,
                                           uint32_t *__restrict__ counts, const uint64_t *__restrict__ offsets,
                                           uint32_t *__restrict__ out) {
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  double p[D];
#pragma unroll
  for (int d = 0; d < D; ++d) p[d] = picks[qi * D + d];
  const uint64_t o = FILL ? offsets[qi] : 0;
  uint32_t cnt = 0;
  if (sp.node_cnt != 0) {
    const uint32_t root = (uint32_t)sp.node_off;
    double rmx[D];
#pragma unroll
    for (int d = 0; d < D; ++d) rmx[d] = tv.nmin[(uint64_t)root * D + d] + tv.nsize[(uint64_t)root * D + d];
    if (box_contains_point<D>(&tv.nmin[(uint64_t)root * D], rmx, p, tol)) {
      int top = 0;
      uint32_t node = root;
      for (;;) {
        if (node_contains_point<D>(&tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], p, tol)) {
          const NodeInfo ni = gload(tv.info + node);
          for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
            bool ok;
            if (BOX) {
              double emn[D], emx[D];
#pragma unroll
              for (int d = 0; d < D; ++d) {
                emn[d] = tv.geom[(uint64_t)d * tv.n + j];
                emx[d] = tv.geom[(uint64_t)(D + d) * tv.n + j];
              }
              ok = box_contains_point<D>(emn, emx, p, tol);
            } else {
              double d2 = 0.0;
#pragma unroll
              for (int d = 0; d < D; ++d) {
                const double x = tv.geom[(uint64_t)d * tv.n + j] - p[d];
                d2 += x * x;
              }
              ok = d2 <= tol * tol;
            }
            if (ok) {
              if (FILL) out[o + cnt] = tv.ids[j];
              ++cnt;
            }
          }
          if (ni.ccount) {
            stack.nx(top) = ni.cbegin + ni.ccount - 1;
            stack.left(top) = (uint8_t)ni.ccount;
            ++top;
          }
        }
        if (top == 0) break;
        node = stack.nx(top - 1)--;
        if (--stack.left(top - 1) == 0) --top;
      }
    }
  }
  if (!FILL) counts[qi] = cnt;
}

template <int D, bool BOX, bool FILL>
__global__ void __launch_bounds__(kQB) k_pick(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                               const double *__restrict__ picks, const uint32_t *__restrict__ qset,
                                               const uint32_t *__restrict__ order, uint64_t nq, double tol,
                                               uint32_t *__restrict__ counts, const uint64_t *__restrict__ offsets,
                                               uint32_t *__restrict__ out) {
  __shared__ uint32_t s_a[kRangeStack * kQB];
  __shared__ uint8_t s_b[kRangeStack * kQB];
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nq) return;
  const uint64_t qi = order ? order[ti] : ti;
  LdsLevelIter S{&s_a[threadIdx.x], &s_b[threadIdx.x]};
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    pick_query<D, BOX, FILL>(tv, qi, picks, qset, tol, S, counts, offsets, out);
  });
}

// ---------------------------------------------------------------------------
// ray first hit
// ---------------------------------------------------------------------------
template <int D>
struct RayTester {
  double origin[D], inv[D], dir[D];
  double min_tol, inc;
  uint32_t sign, zero;  // bit masks

  // internal_geometry_module.h:539-568 Make
  __device__ bool make(const double *o, const double *dv, double tol, double tinc) {
    double s2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) s2 += dv[d] * dv[d];
    if (!(fabs(s2 - 1.0) < 0.000001)) return false;
    min_tol = tol;
    inc = tinc;
    sign = 0;
    zero = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      origin[d] = o[d];
      dir[d] = dv[d];
      const bool z = fabs(dv[d]) < 2.0 * DBL_EPSILON;
      inv[d] = z ? __builtin_nan("") : (1.0 / dv[d]);
      if (dv[d] < 0) sign |= 1u << d;
      if (z) zero |= 1u << d;
    }
    return true;
  }

  // BoxHitTest<isConeToleranceConsidered = true> igm:631-712
  __device__ bool box_hit(const double (&mind)[D], const double (&maxd)[D], double *enter, double *exit) const {
    double en = -DBL_MAX, ex = DBL_MAX;
    double pd0[D], pd1[D];
#pragma unroll
    for (int d = 0; d < D; ++d) pd0[d] = mind[d] * inv[d];
#pragma unroll
    for (int d = 0; d < D; ++d) pd1[d] = maxd[d] * inv[d];
    if (zero) {
#pragma unroll
      for (int d = 0; d < D; ++d)
        if (!((zero >> d) & 1u)) en = fmax(en, ((sign >> d) & 1u) ? pd1[d] : pd0[d]);
#pragma unroll
      for (int d = 0; d < D; ++d)
        if (!((zero >> d) & 1u)) ex = fmin(ex, ((sign >> d) & 1u) ? pd0[d] : pd1[d]);
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d) en = smax(en, ((sign >> d) & 1u) ? pd1[d] : pd0[d]);
#pragma unroll
      for (int d = 0; d < D; ++d) ex = smin(ex, ((sign >> d) & 1u) ? pd0[d] : pd1[d]);
    }
    const double exit_tol = smax(0.0, ex) * inc + min_tol;
    en -= 0.5 * (smax(0.0, en) * inc + min_tol);
    ex += 0.5 * (smax(0.0, ex) * inc + min_tol);
    if (en >= ex || ex <= 0.0) return false;
    en = smax(0.0, en);
    if (zero) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (!((zero >> d) & 1u)) continue;
        if (mind[d] > exit_tol || maxd[d] < -exit_tol) return false;
      }
    }
    *enter = en;
    *exit = ex;
    return true;
  }

  // Hit(minPoint, size) igm:575-585
  __device__ bool hit_node(const double *mn, const double *sz, double *en, double *ex) const {
    double a[D], b[D];
#pragma unroll
    for (int d = 0; d < D; ++d) a[d] = mn[d] - origin[d];
#pragma unroll
    for (int d = 0; d < D; ++d) b[d] = mn[d] + sz[d] - origin[d];
    return box_hit(a, b, en, ex);
  }

  // Hit(point) igm:599-626
  __device__ bool hit_point(const double (&p)[D], double *en, double *ex) const {
    double v[D];
#pragma unroll
    for (int d = 0; d < D; ++d) v[d] = p[d] - origin[d];
    double t = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) t += v[d] * dir[d];
    double len2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) len2 += v[d] * v[d];
    const double dist2 = smax(0.0, len2 - t * t);
    const double ct = smax(0.0, t) * inc + min_tol;
    if (dist2 <= ct * ct && t >= -ct) {
      *en = smax(0.0, t - ct);
      *ex = t + ct;
      return true;
    }
    return false;
  }
};

// Descending enter distance, ties by id (selection into output order).
template <int RCAP>
__device__ __forceinline__ void ray_output(uint64_t qi, uint32_t cnt, double (&ce)[RCAP], uint32_t (&cid)[RCAP],
                                           uint32_t *__restrict__ out_ids, uint32_t *__restrict__ out_cnt) {
  out_cnt[qi] = cnt;
  for (uint32_t a = 0; a < cnt; ++a) {
    uint32_t best = a;
    for (uint32_t b = a + 1; b < cnt; ++b) {
      double eb = 0, ebest = 0;
      uint32_t ib = 0, ibest = 0;
#pragma unroll
      for (int t = 0; t < RCAP; ++t) {
        if ((uint32_t)t == b) { eb = ce[t]; ib = cid[t]; }
        if ((uint32_t)t == best) { ebest = ce[t]; ibest = cid[t]; }
      }
      if (eb > ebest || (eb == ebest && ib < ibest)) best = b;
    }
    double ea = 0, ebest = 0;
    uint32_t ia = 0, ibest = 0;
#pragma unroll
    for (int t = 0; t < RCAP; ++t) {
      if ((uint32_t)t == a) { ea = ce[t]; ia = cid[t]; }
      if ((uint32_t)t == best) { ebest = ce[t]; ibest = cid[t]; }
    }
#pragma unroll
    for (int t = 0; t < RCAP; ++t) {
      if ((uint32_t)t == a) { ce[t] = ebest; cid[t] = ibest; }
      if ((uint32_t)t == best && best != a) { ce[t] = ea; cid[t] = ia; }
    }
    out_ids[qi * RCAP + a] = ibest;
  }
}

// Node heaps for the ray walk: a binary min-heap on node enter distance; push
// sifts up while smaller than the parent, pop moves the last entry to the root and
// sifts it down towards the smaller child, so equal enter distances leave in a
// fixed order.  LdsHeap: lane-interleaved in LDS; ArrayHeap: private.
constexpr int kRayHeap = 16;  // 16 x 12 B x 64 lanes = 12 KiB per block

struct LdsHeap {
  static constexpr int cap = kRayHeap;
  uint32_t *n;
  double *p;
  __device__ __forceinline__ int capacity() const { return cap; }
  __device__ __forceinline__ uint32_t &node(int i) { return n[i * kQB]; }
  __device__ __forceinline__ double &prio(int i) { return p[i * kQB]; }
};

template <int D>
struct ArrayHeap {
  static constexpr int cap = 2 * StackCap<D>::value;
  uint32_t nd[cap];
  double pr[cap];
  __device__ __forceinline__ int capacity() const { return cap; }
  __device__ __forceinline__ uint32_t &node(int i) { return nd[i]; }
  __device__ __forceinline__ double &prio(int i) { return pr[i]; }
};

// Per-ray heap in global memory holding up to `c` nodes: with c = the tree's node
// count it never fills (a node is queued at most once).  For the rare rays whose
// frontier outgrows the private heap.
struct GlobalHeap {
  uint32_t *nd;
  double *pr;
  int c;
  __device__ __forceinline__ int capacity() const { return c; }
  __device__ __forceinline__ uint32_t &node(int i) { return nd[i]; }
  __device__ __forceinline__ double &prio(int i) { return pr[i]; }
};

template <typename Heap>
__device__ __forceinline__ void heap_push(Heap &H, int &n, uint32_t node, double prio) {
  int i = n++;
  H.node(i) = node;
  H.prio(i) = prio;
  while (i > 0) {
    const int par = (i - 1) / 2;
    if (!(H.prio(i) < H.prio(par))) break;
    const uint32_t tn = H.node(par);
    const double tp = H.prio(par);
    H.node(par) = H.node(i);
    H.prio(par) = H.prio(i);
    H.node(i) = tn;
    H.prio(i) = tp;
    i = par;
  }
}

template <typename Heap>
__device__ __forceinline__ void heap_pop(Heap &H, int &n, uint32_t *node, double *prio) {
  *node = H.node(0);
  *prio = H.prio(0);
  --n;
  H.node(0) = H.node(n);
  H.prio(0) = H.prio(n);
  int i = 0;
  for (;;) {
    const int l = 2 * i + 1, r = l + 1;
    int m = i;
    if (l < n && H.prio(l) < H.prio(m)) m = l;
    if (r < n && H.prio(r) < H.prio(m)) m = r;
    if (m == i) break;
    const uint32_t tn = H.node(m);
    const double tp = H.prio(m);
    H.node(m) = H.node(i);
    H.prio(m) = H.prio(i);
    H.node(i) = tn;
    H.prio(i) = tp;
    i = m;
  }
}

// RayIntersectedFirst (ot_query.h:2035-2122): best-first on node enter distance
// (TraverseNodesByPriority :235-291), stopping at the first node entering beyond
// the bound M; every hit child is queued (M may still rise before it is popped).
// Each entity hit entering within M is a candidate; a hit whose exit is below M
// sets M = (1 + inc) exit + tol, which can raise or lower it, so the result depends
// on this visiting order -- the reference's.  Returns false (nothing written) when
// the heap would overflow.
template <int D, bool BOX, int RCAP, typename Heap>
__device__ __forceinline__ bool ray_query(const TreeView &tv, uint64_t qi, const double *__restrict__ rays,
                                          const uint32_t *__restrict__ qset, double tol, double inc, double max_dist,
                                          Heap &H, uint32_t *__restrict__ out_ids, uint32_t *__restrict__ out_cnt,
                                          uint32_t *__restrict__ overflow, uint32_t *__restrict__ band,
                                          double *__restrict__ out_M) {
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  uint32_t cnt = 0;
  double ce[RCAP];
  uint32_t cid[RCAP];
  RayTester<D> rt;
  double M = max_dist;
  bool over = false;  // more candidates than the list: redone by k_ray_band
  if (sp.node_cnt != 0 && rt.make(&rays[qi * 2 * D], &rays[qi * 2 * D + D], tol, inc)) {
    const double applied = (tol == 0) ? DBL_EPSILON : tol;
    int n = 0;
    double en, ex;
    const uint32_t root = (uint32_t)sp.node_off;
    if (rt.hit_node(&tv.nmin[(uint64_t)root * D], &tv.nsize[(uint64_t)root * D], &en, &ex)) heap_push(H, n, root, en);
    while (n > 0) {
      uint32_t node;
      double prio;
      heap_pop(H, n, &node, &prio);
      if (prio > M) break;
      const NodeInfo ni = gload(tv.info + node);
      for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
        bool hit;
        if (BOX) {
          double lo[D], hi[D];
#pragma unroll
          for (int d = 0; d < D; ++d) lo[d] = tv.geom[(uint64_t)d * tv.n + j] - rt.origin[d];
#pragma unroll
          for (int d = 0; d < D; ++d) hi[d] = tv.geom[(uint64_t)(D + d) * tv.n + j] - rt.origin[d];
          hit = rt.box_hit(lo, hi, &en, &ex);
        } else {
          double pnt[D];
#pragma unroll
          for (int d = 0; d < D; ++d) pnt[d] = tv.geom[(uint64_t)d * tv.n + j];
          hit = rt.hit_point(pnt, &en, &ex);
        }
        if (!hit || en > M) continue;
        if (cnt == RCAP) {
          over = true;
        } else {
#pragma unroll
          for (int t = 0; t < RCAP; ++t)
            if ((uint32_t)t == cnt) {
              ce[t] = en;
              cid[t] = tv.ids[j];
            }
          ++cnt;
        }
        if (ex >= M) continue;
        M = (1.0 + inc) * ex + applied;
        uint32_t w = 0;
#pragma unroll
        for (int t = 0; t < RCAP; ++t) {
          if ((uint32_t)t < cnt && !(ce[t] > M)) {
            // compact in place (w <= t)
#pragma unroll
            for (int u = 0; u < RCAP; ++u)
              if ((uint32_t)u == w) {
                ce[u] = ce[t];
                cid[u] = cid[t];
              }
            ++w;
          }
        }
        cnt = w;
      }
      for (uint32_t c = 0; c < ni.ccount; ++c) {
        const uint32_t x = ni.cbegin + c;
        if (!rt.hit_node(&tv.nmin[(uint64_t)x * D], &tv.nsize[(uint64_t)x * D], &en, &ex)) continue;
        if (n == H.capacity()) return false;
        heap_push(H, n, x, en);
      }
    }
  }
  if (over) {
    band[atomicAdd(overflow, 1u)] = (uint32_t)qi;
    out_M[qi] = M;
  }
  ray_output<RCAP>(qi, cnt, ce, cid, out_ids, out_cnt);
  return true;
}

template <int D, bool BOX, int RCAP>
__global__ void __launch_bounds__(kQB) k_ray(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                              const double *__restrict__ rays, const uint32_t *__restrict__ qset,
                                              const uint32_t *__restrict__ order, uint64_t nq, double tol, double inc,
                                              double max_dist, uint32_t *__restrict__ out_ids,
                                              uint32_t *__restrict__ out_cnt, uint32_t *__restrict__ overflow,
                                              uint32_t *__restrict__ band, double *__restrict__ out_M,
                                              uint32_t *__restrict__ redo, uint32_t *__restrict__ redo_n) {
  __shared__ uint32_t s_n[kRayHeap * kQB];
  __shared__ double s_p[kRayHeap * kQB];
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nq) return;
  const uint64_t qi = order ? order[ti] : ti;
  LdsHeap H{&s_n[threadIdx.x], &s_p[threadIdx.x]};
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    if (!ray_query<D, BOX, RCAP>(tv, qi, rays, qset, tol, inc, max_dist, H, out_ids, out_cnt, overflow, band, out_M))
      redo[atomicAdd(redo_n, 1u)] = (uint32_t)ti;  // position in the query order
  });
}

// RayIntersectedFirst with one wavefront per ray, for batches too small to fill the GPU
// one lane per ray.  The walk is ray_query's: the same node heap (run by lane 0 in LDS,
// kRayWaveHeap nodes; a ray that fills it is listed for the deep replay like k_ray's),
// the same stop at the first node entering beyond M, a node's entities tested 64 at a
// time and their hits taken in index order against the M of that moment, the same M
// updates and list compaction, hit children queued in index order.  The candidate list
// is spread over the lanes (64 slots); a ray left with more than RCAP candidates, or
// one that outgrows 64, goes to the band replay as k_ray's overflowing rays do.
constexpr int kRayWaveHeap = 256;
// Measured on 4M loose boxes (C2 tree), random rays, lane walk vs wavefront walk: 10k
// rays 2.74 vs 0.42 ms, 100k 4.36 vs 1.61, 400k 8.26 vs 5.38, 1.6M 21.9 vs 20.0; the
// lane walk is kept for larger batches, where the two meet.
constexpr uint64_t kRayWaveMaxQueries = 1ull << 22;

template <int D, bool BOX, int RCAP>
__global__ void __launch_bounds__(64 * kWaveQueries)
    k_ray_wave(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview, const double *__restrict__ rays,
               const uint32_t *__restrict__ qset, const uint32_t *__restrict__ order, uint64_t nq, double tol, double inc,
               double max_dist, uint32_t *__restrict__ out_ids, uint32_t *__restrict__ out_cnt,
               uint32_t *__restrict__ overflow, uint32_t *__restrict__ band, double *__restrict__ out_M,
               uint32_t *__restrict__ redo, uint32_t *__restrict__ redo_n) {
  static_assert(RCAP <= 64, "the output rows fit the lanes");
  __shared__ uint32_t s_hn[kWaveQueries][kRayWaveHeap];
  __shared__ double s_hp[kWaveQueries][kRayWaveHeap];
  __shared__ double s_ce[kWaveQueries][64];
  __shared__ uint32_t s_cid[kWaveQueries][64];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t ti = (uint64_t)blockIdx.x * kWaveQueries + wv;
  if (ti >= nq) return;  // wavefront-uniform
  const uint64_t qi = order ? order[ti] : ti;
  const TreeView tv = views[qview ? qview[qi] : 0u];
  const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
  struct WaveHeap {
    uint32_t *nd;
    double *pr;
    __device__ __forceinline__ int capacity() const { return kRayWaveHeap; }
    __device__ __forceinline__ uint32_t &node(int i) { return nd[i]; }
    __device__ __forceinline__ double &prio(int i) { return pr[i]; }
  } H{s_hn[wv], s_hp[wv]};
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t cnt = 0;
  double ce = 0.0;
  uint32_t cid = 0;
  RayTester<D> rt;
  double M = max_dist;
  bool over = false, heap_full = false;
  if (sp.node_cnt != 0 && rt.make(&rays[qi * 2 * D], &rays[qi * 2 * D + D], tol, inc)) {
    const double applied = (tol == 0) ? DBL_EPSILON : tol;
    int n = 0;
    double en, ex;
    const uint32_t root = (uint32_t)sp.node_off;
    if (rt.hit_node(&tv.nmin[(uint64_t)root * D], &tv.nsize[(uint64_t)root * D], &en, &ex)) {
      if (lane == 0) heap_push(H, n, root, en);
      n = 1;
    }
    while (n > 0) {
      uint32_t node = 0;
      double prio = 0.0;
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) heap_pop(H, n, &node, &prio);
      __builtin_amdgcn_wave_barrier();
      n = __shfl(n, 0, 64);
      node = __shfl(node, 0, 64);
      prio = __shfl(prio, 0, 64);
      if (prio > M) break;
      const NodeInfo ni = gload(tv.info + node);
      for (uint32_t base = ni.ebegin; base < ni.ebegin + ni.ecount; base += 64) {
        const uint32_t j = base + lane;
        const bool valid = j < ni.ebegin + ni.ecount;
        bool hit = false;
        double jen = 0.0, jex = 0.0;
        uint32_t jid = 0;
        if (valid) {
          jid = tv.ids[j];
          if (BOX) {
            double lo[D], hi[D];
#pragma unroll
            for (int d = 0; d < D; ++d) lo[d] = tv.geom[(uint64_t)d * tv.n + j] - rt.origin[d];
#pragma unroll
            for (int d = 0; d < D; ++d) hi[d] = tv.geom[(uint64_t)(D + d) * tv.n + j] - rt.origin[d];
            hit = rt.box_hit(lo, hi, &jen, &jex);
          } else {
            double pnt[D];
#pragma unroll
            for (int d = 0; d < D; ++d) pnt[d] = tv.geom[(uint64_t)d * tv.n + j];
            hit = rt.hit_point(pnt, &jen, &jex);
          }
        }
        uint64_t mask = __ballot(hit);
        while (mask) {
          const int x = __builtin_ctzll(mask);
          mask &= mask - 1;
          const double hen = __shfl(jen, x, 64), hex = __shfl(jex, x, 64);
          const uint32_t hid = __shfl(jid, x, 64);
          if (hen > M) continue;
          if (cnt == 64u) {
            over = true;
          } else {
            if (lane == cnt) {
              ce = hen;
              cid = hid;
            }
            ++cnt;
          }
          if (hex >= M) continue;
          M = (1.0 + inc) * hex + applied;
          // keep the candidates entering within the new M, in their order
          const bool keep = lane < cnt && !(ce > M);
          const uint64_t km = __ballot(keep);
          const uint32_t kept = (uint32_t)__popcll(km);
          if (kept != cnt) {
            if (keep) {
              const uint32_t at = (uint32_t)__popcll(km & lt);
              s_ce[wv][at] = ce;
              s_cid[wv][at] = cid;
            }
            __builtin_amdgcn_wave_barrier();
            cnt = kept;
            if (lane < cnt) {
              ce = s_ce[wv][lane];
              cid = s_cid[wv][lane];
            }
            __builtin_amdgcn_wave_barrier();
          }
        }
      }
      if (ni.ccount) {
        const uint32_t x = ni.cbegin + lane;
        double cen = 0.0, cex = 0.0;
        const bool chit = lane < ni.ccount && rt.hit_node(&tv.nmin[(uint64_t)x * D], &tv.nsize[(uint64_t)x * D], &cen, &cex);
        uint64_t cm = __ballot(chit);
        const int pushes = __popcll(cm);
        if (n + pushes > kRayWaveHeap) {
          heap_full = true;
          break;
        }
        __builtin_amdgcn_wave_barrier();
        while (cm) {
          const int c = __builtin_ctzll(cm);
          cm &= cm - 1;
          const double p = __shfl(cen, c, 64);
          if (lane == 0) heap_push(H, n, ni.cbegin + (uint32_t)c, p);
        }
        __builtin_amdgcn_wave_barrier();
        n = __shfl(n, 0, 64);
      }
    }
  }
  if (heap_full) {  // replayed by the deep walk with a larger heap
    if (lane == 0) redo[atomicAdd(redo_n, 1u)] = (uint32_t)ti;
    return;
  }
  if (cnt > (uint32_t)RCAP) over = true;
  if (lane == 0 && over) {
    band[atomicAdd(overflow, 1u)] = (uint32_t)qi;
    out_M[qi] = M;
  }
  // descending enter distance, ties by id: bitonic sort over the lanes (empty slots last)
  double key = lane < cnt ? ce : -__builtin_inf();
  uint32_t kid = lane < cnt ? cid : 0xFFFFFFFFu;
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const double ok = __shfl_xor(key, stride, 64);
      const uint32_t oi = __shfl_xor(kid, stride, 64);
      const bool lower = (lane & stride) == 0;
      const bool up = (lane & size) == 0;  // this block sorts first-before-second
      // "first" = larger enter, then smaller id
      const bool other_first = ok > key || (ok == key && oi < kid);
      const bool take = (lower == up) ? other_first : !other_first && !(ok == key && oi == kid);
      if (take) {
        key = ok;
        kid = oi;
      }
    }
  }
  const uint32_t nout = cnt < (uint32_t)RCAP ? cnt : (uint32_t)RCAP;
  if (lane == 0) out_cnt[qi] = nout;
  if (lane < nout) out_ids[qi * RCAP + lane] = kid;
}

// Rays whose LDS heap would overflow, with the large private heap (GLOBAL: with a
// heap of hcap nodes in global memory, heap_n / heap_p + ti * hcap).  A ray that
// outgrows the private heap is listed in fail_list (count in *fail).
template <int D, bool BOX, int RCAP, bool GLOBAL>
__global__ void __launch_bounds__(kQB) k_ray_deep(const TreeView *__restrict__ views,
                                                   const uint32_t *__restrict__ qview, const double *__restrict__ rays,
                                                   const uint32_t *__restrict__ qset, const uint32_t *__restrict__ order,
                                                   const uint32_t *__restrict__ list, uint32_t nlist, double tol,
                                                   double inc, double max_dist, uint32_t *__restrict__ out_ids,
                                                   uint32_t *__restrict__ out_cnt, uint32_t *__restrict__ overflow,
                                                   uint32_t *__restrict__ band, double *__restrict__ out_M,
                                                   uint32_t *__restrict__ fail, uint32_t *__restrict__ fail_list,
                                                   uint32_t *__restrict__ heap_n, double *__restrict__ heap_p,
                                                   uint32_t hcap) {
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nlist) return;
  const uint64_t qi = order ? order[list[ti]] : list[ti];
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    bool ok;
    if constexpr (GLOBAL) {
      GlobalHeap H{heap_n + ti * hcap, heap_p + ti * hcap, (int)hcap};
      ok = ray_query<D, BOX, RCAP>(tv, qi, rays, qset, tol, inc, max_dist, H, out_ids, out_cnt, overflow, band, out_M);
    } else {
      ArrayHeap<D> H;
      ok = ray_query<D, BOX, RCAP>(tv, qi, rays, qset, tol, inc, max_dist, H, out_ids, out_cnt, overflow, band, out_M);
    }
    if (!ok) fail_list[atomicAdd(fail, 1u)] = list[ti];
  });
}

constexpr uint32_t kReplayListFull = 0xFFFFFFFFu;  // candidate list outgrown: replay with a longer one
constexpr uint32_t kReplayHeapFull = 0xFFFFFFFEu;  // node heap outgrown: replay with a global heap

// Rays with more candidates than the register list are replayed exactly (same
// best-first walk, same bound updates) with a candidate list of `cap` entries in
// global memory; a replay that outgrows it sets its count to UINT32_MAX and is run
// again with a larger list.  The candidates end sorted by descending enter, ties by id.
// GLOBAL: the node heap in global memory (heap_n / heap_p + ti * hcap), for rays
// whose frontier outgrows the private heap; a ray that finds its heap full is
// listed as failed (counts[ti] = UINT32_MAX - 1, *fail set) and replayed with it.
template <int D, bool BOX, bool GLOBAL>
__global__ void __launch_bounds__(kQB) k_ray_replay(const TreeView *__restrict__ views,
                                                     const uint32_t *__restrict__ qview,
                                                     const double *__restrict__ rays, const uint32_t *__restrict__ qset,
                                                     const uint32_t *__restrict__ list, uint32_t nlist, double tol,
                                                     double inc, double max_dist, uint32_t cap,
                                                     double *__restrict__ cand_e, uint32_t *__restrict__ cand_id,
                                                     uint32_t *__restrict__ counts, uint32_t *__restrict__ fail,
                                                     uint32_t *__restrict__ heap_n, double *__restrict__ heap_p,
                                                     uint32_t hcap) {
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nlist) return;
  const uint64_t qi = list[ti];
  double *ce = cand_e + ti * cap;
  uint32_t *cid = cand_id + ti * cap;
  using Heap = std::conditional_t<GLOBAL, GlobalHeap, ArrayHeap<D>>;
  Heap H;
  if constexpr (GLOBAL) H = GlobalHeap{heap_n + ti * hcap, heap_p + ti * hcap, (int)hcap};
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
    uint32_t cnt = 0;
    bool too_many = false, heap_full = false;
    RayTester<D> rt;
    if (sp.node_cnt != 0 && rt.make(&rays[qi * 2 * D], &rays[qi * 2 * D + D], tol, inc)) {
      const double applied = (tol == 0) ? DBL_EPSILON : tol;
      double M = max_dist;
      int n = 0;
      double en, ex;
      const uint32_t root = (uint32_t)sp.node_off;
      if (rt.hit_node(&tv.nmin[(uint64_t)root * D], &tv.nsize[(uint64_t)root * D], &en, &ex)) heap_push(H, n, root, en);
      while (n > 0 && !too_many) {
        uint32_t node;
        double prio;
        heap_pop(H, n, &node, &prio);
        if (prio > M) break;
        const NodeInfo ni = gload(tv.info + node);
        for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
          bool hit;
          if (BOX) {
            double lo[D], hi[D];
#pragma unroll
            for (int d = 0; d < D; ++d) lo[d] = tv.geom[(uint64_t)d * tv.n + j] - rt.origin[d];
#pragma unroll
            for (int d = 0; d < D; ++d) hi[d] = tv.geom[(uint64_t)(D + d) * tv.n + j] - rt.origin[d];
            hit = rt.box_hit(lo, hi, &en, &ex);
          } else {
            double pnt[D];
#pragma unroll
            for (int d = 0; d < D; ++d) pnt[d] = tv.geom[(uint64_t)d * tv.n + j];
            hit = rt.hit_point(pnt, &en, &ex);
          }
          if (!hit || en > M) continue;
          if (cnt == cap) {
            too_many = true;
            break;
          }
          // candidates form a max-heap on enter distance (the reference's priority
          // queue): a lower bound pops the entries above it from the top
          {
            uint32_t i = cnt++;
            while (i > 0) {
              const uint32_t par = (i - 1) / 2;
              if (!(ce[par] < en)) break;
              ce[i] = ce[par];
              cid[i] = cid[par];
              i = par;
            }
            ce[i] = en;
            cid[i] = tv.ids[j];
          }
          if (ex >= M) continue;
          M = (1.0 + inc) * ex + applied;
          while (cnt > 0 && ce[0] > M) {
            --cnt;
            const double le = ce[cnt];
            const uint32_t li = cid[cnt];
            uint32_t i = 0;
            for (;;) {
              const uint32_t l = 2 * i + 1, r = l + 1;
              uint32_t m = i;
              double vm = le;
              if (l < cnt && vm < ce[l]) m = l, vm = ce[l];
              if (r < cnt && vm < ce[r]) m = r;
              if (m == i) break;
              ce[i] = ce[m];
              cid[i] = cid[m];
              i = m;
            }
            ce[i] = le;
            cid[i] = li;
          }
        }
        for (uint32_t c = 0; c < ni.ccount && !too_many; ++c) {
          const uint32_t x = ni.cbegin + c;
          if (!rt.hit_node(&tv.nmin[(uint64_t)x * D], &tv.nsize[(uint64_t)x * D], &en, &ex)) continue;
          if (n == H.capacity()) {
            atomicOr(fail, 1u);
            heap_full = true;
            too_many = true;
            break;
          }
          heap_push(H, n, x, en);
        }
      }
    }
    if (too_many) {
      counts[ti] = heap_full ? kReplayHeapFull : kReplayListFull;
      return;
    }
    counts[ti] = cnt;
    // heap sort: descending enter, ties by id
    auto sift = [&](uint32_t i, uint32_t m) {
      for (;;) {
        uint32_t l = 2 * i + 1, r = l + 1, x = i;
        if (l < m && pair_less<true>(ce[x], cid[x], ce[l], cid[l])) x = l;
        if (r < m && pair_less<true>(ce[x], cid[x], ce[r], cid[r])) x = r;
        if (x == i) return;
        const double td = ce[i];
        ce[i] = ce[x];
        ce[x] = td;
        const uint32_t tid = cid[i];
        cid[i] = cid[x];
        cid[x] = tid;
        i = x;
      }
    };
    for (uint32_t i = cnt / 2; i-- > 0;) sift(i, cnt);
    for (uint32_t m = cnt; m > 1; --m) {
      const double td = ce[0];
      ce[0] = ce[m - 1];
      ce[m - 1] = td;
      const uint32_t tid = cid[0];
      cid[0] = cid[m - 1];
      cid[m - 1] = tid;
      sift(0, m - 1);
    }
  });
}

// Copies replayed candidate lists (packed, list i at src + src_off[i]) into the CSR
// slots of their rays.
__global__ void k_replay_splice(const uint32_t *__restrict__ rays, uint32_t nrays, const uint64_t *__restrict__ src_off,
                                const uint32_t *__restrict__ src, const uint64_t *__restrict__ off,
                                uint32_t *__restrict__ dst) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrays) return;
  const uint64_t o = off[rays[i]];
  for (uint64_t t = src_off[i]; t < src_off[i + 1]; ++t) dst[o + (t - src_off[i])] = src[t];
}

// ---------------------------------------------------------------------------
// ray, all hits
// ---------------------------------------------------------------------------
// RayIntersectedAll (ot_query.h:1929-2012): depth-first (last child first)
// skipping nodes the ray misses; every entity hit whose enter distance is within
// max_dist (any, when max_dist is 0) is kept.  Count pass, then a fill pass that
// writes (enter distance, id) at the query's CSR offset; k_sort_band orders each
// list ascending by enter distance (ties by id).
template <int D, bool BOX, bool FILL>
__device__ __forceinline__ void ray_all_query(const TreeView &tv, uint64_t qi, const double *__restrict__ rays,
                                              const uint32_t *__restrict__ qset, double tol, double inc,
                                              double max_dist, LdsLevelIter &stack, uint32_t *__restrict__ counts,
                                              const uint64_t *__restrict__ offsets, double *__restrict__ out_d,
                                              uint32_t *__restrict__ out_id) {
  const uint32_t s = qset ? qset[qi] : 0u;
  const SetParams &sp = tv.sets[s];
  const uint64_t o = FILL ? offsets[qi] : 0;
  uint32_t cnt = 0;
  RayTester<D> rt;
  if (sp.node_cnt != 0 && rt.make(&rays[qi * 2 * D], &rays[qi * 2 * D + D], tol, inc)) {
    int top = 0;
    uint32_t node = (uint32_t)sp.node_off;
    double en, ex;
    for (;;) {
      if (rt.hit_node(&tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &en, &ex)) {
        const NodeInfo ni = gload(tv.info + node);
        for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
          bool hit;
          if (BOX) {
            double lo[D], hi[D];
#pragma unroll
            for (int d = 0; d < D; ++d) lo[d] = tv.geom[(uint64_t)d * tv.n + j] - rt.origin[d];
#pragma unroll
            for (int d = 0; d < D; ++d) hi[d] = tv.geom[(uint64_t)(D + d) * tv.n + j] - rt.origin[d];
            hit = rt.box_hit(lo, hi, &en, &ex);
          } else {
            double pnt[D];
#pragma unroll
            for (int d = 0; d < D; ++d) pnt[d] = tv.geom[(uint64_t)d * tv.n + j];
            hit = rt.hit_point(pnt, &en, &ex);
          }
          if (!hit || !(max_dist == 0 || en <= max_dist)) continue;
          if (FILL) {
            out_d[o + cnt] = en;
            out_id[o + cnt] = tv.ids[j];
          }
          ++cnt;
        }
        if (ni.ccount) {
          stack.nx(top) = ni.cbegin + ni.ccount - 1;
          stack.left(top) = (uint8_t)ni.ccount;
          ++top;
        }
      }
      if (top == 0) break;
      node = stack.nx(top - 1)--;
      if (--stack.left(top - 1) == 0) --top;
    }
  }
  if (!FILL) counts[qi] = cnt;
}

template <int D, bool BOX, bool FILL>
__global__ void __launch_bounds__(kQB) k_ray_all(const TreeView *__restrict__ views, const uint32_t *__restrict__ qview,
                                                  const double *__restrict__ rays, const uint32_t *__restrict__ qset,
                                                  const uint32_t *__restrict__ order, uint64_t nq, double tol,
                                                  double inc, double max_dist, uint32_t *__restrict__ counts,
                                                  const uint64_t *__restrict__ offsets, double *__restrict__ out_d,
                                                  uint32_t *__restrict__ out_id) {
  __shared__ uint32_t s_a[kRangeStack * kQB];
  __shared__ uint8_t s_b[kRangeStack * kQB];
  const uint64_t ti = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (ti >= nq) return;
  const uint64_t qi = order ? order[ti] : ti;
  LdsLevelIter S{&s_a[threadIdx.x], &s_b[threadIdx.x]};
  with_tree(views, qview, qi, [&](const TreeView &tv) {
    ray_all_query<D, BOX, FILL>(tv, qi, rays, qset, tol, inc, max_dist, S, counts, offsets, out_d, out_id);
  });
}

// RayIntersectedAll with one wavefront per ray (2-D / 3-D), for batches too small to
// fill the GPU one lane per ray: the frontier expansion and per-lane subtree walks of
// k_range_wave, visiting the nodes the ray hits and keeping every entity hit within
// max_dist, appended at the ray's offset in no particular order (the lists are sorted
// by enter distance, ties by id, afterwards).  The same hits as ray_all_query's walk.
template <int D, bool BOX, bool FILL>
__global__ void __launch_bounds__(64) k_ray_all_wave(const TreeView *__restrict__ views,
                                                      const uint32_t *__restrict__ qview,
                                                      const double *__restrict__ rays, const uint32_t *__restrict__ qset,
                                                      const uint32_t *__restrict__ order, uint64_t nq, double tol,
                                                      double inc, double max_dist, uint32_t *__restrict__ counts,
                                                      const uint64_t *__restrict__ offsets, double *__restrict__ out_d,
                                                      uint32_t *__restrict__ out_id) {
  static_assert(D <= 3, "frontier sized for up to 8 children per node");
  __shared__ uint32_t s_front[64 * 8];
  __shared__ uint32_t s_nf[2], s_cnt, s_next;
  __shared__ uint32_t s_a[kRangeStack * 64];
  __shared__ uint8_t s_b[kRangeStack * 64];
  const uint64_t ti = blockIdx.x;
  if (ti >= nq) return;
  const uint32_t lane = threadIdx.x;
  const uint64_t qi = order ? order[ti] : ti;
  const TreeView &tv = views[qview ? qview[qi] : 0u];
  const SetParams &sp = tv.sets[qset ? qset[qi] : 0u];
  RayTester<D> rt;
  const bool walk = sp.node_cnt != 0 && rt.make(&rays[qi * 2 * D], &rays[qi * 2 * D + D], tol, inc);
  const uint64_t o = FILL ? offsets[qi] : 0;
  // node visit: the node's entities hit, and whether to go on into its children
  auto visit = [&](uint32_t node, NodeInfo &ni) -> bool {
    double en, ex;
    if (!rt.hit_node(&tv.nmin[(uint64_t)node * D], &tv.nsize[(uint64_t)node * D], &en, &ex)) return false;
    ni = gload(tv.info + node);
    for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
      bool hit;
      if (BOX) {
        double lo[D], hi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) lo[d] = tv.geom[(uint64_t)d * tv.n + j] - rt.origin[d];
#pragma unroll
        for (int d = 0; d < D; ++d) hi[d] = tv.geom[(uint64_t)(D + d) * tv.n + j] - rt.origin[d];
        hit = rt.box_hit(lo, hi, &en, &ex);
      } else {
        double pnt[D];
#pragma unroll
        for (int d = 0; d < D; ++d) pnt[d] = tv.geom[(uint64_t)d * tv.n + j];
        hit = rt.hit_point(pnt, &en, &ex);
      }
      if (!hit || !(max_dist == 0 || en <= max_dist)) continue;
      const uint32_t at = atomicAdd(&s_cnt, 1u);
      if (FILL) {
        out_d[o + at] = en;
        out_id[o + at] = tv.ids[j];
      }
    }
    return ni.ccount != 0;
  };
  if (lane == 0) {
    s_cnt = 0;
    s_next = 0;
    s_nf[0] = 0;
    s_nf[1] = 0;
    if (walk) {
      s_front[0] = (uint32_t)sp.node_off;
      s_nf[0] = 1;
    }
  }
  __syncthreads();
  int cur = 0;
  while (s_nf[cur] != 0 && s_nf[cur] < 64) {
    const uint32_t nf = s_nf[cur];
    const uint32_t mine = lane < nf ? s_front[lane] : 0u;  // read before the buffer is rewritten
    if (lane == 0) s_nf[cur ^ 1] = 0;
    __syncthreads();
    if (lane < nf) {
      NodeInfo ni;
      if (visit(mine, ni)) {
        const uint32_t at = atomicAdd(&s_nf[cur ^ 1], ni.ccount);
        for (uint32_t c = 0; c < ni.ccount; ++c) s_front[at + c] = ni.cbegin + c;
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  const uint32_t nf = s_nf[cur];
  LdsLevelIter stack{&s_a[lane], &s_b[lane], 64};
  for (;;) {
    const uint32_t item = atomicAdd(&s_next, 1u);
    if (item >= nf) break;
    int top = 0;
    uint32_t node = s_front[item];
    for (;;) {
      NodeInfo ni;
      if (visit(node, ni)) {
        stack.nx(top) = ni.cbegin + ni.ccount - 1;
        stack.left(top) = (uint8_t)ni.ccount;
        ++top;
      }
      if (top == 0) break;
      node = stack.nx(top - 1)--;
      if (--stack.left(top - 1) == 0) --top;
    }
  }
  __syncthreads();
  if (!FILL && lane == 0) counts[qi] = s_cnt;
}

// ---------------------------------------------------------------------------
// collision (boxes): every entity of the left set queries the right set's tree
// with its own box.  Self collision (SELF: the same set on both sides) keeps
// partners with a larger id, so each unordered pair is found once; two trees
// keep every (left, right) pair.  Pair predicate: GetBoxRelation == Overlapped
// (adapters/general.h:262-289); node pruning: AreBoxesOverlappingByMinPoint on
// the MBR (internal_geometry_module.h:148-161).
// ---------------------------------------------------------------------------
template <int D, bool FILL, bool SELF>
__global__ void __launch_bounds__(kQB) k_collide(TreeView lv, uint32_t lset, TreeView rv, uint32_t rset, double tol,
                                                  uint32_t *__restrict__ counts, const uint64_t *__restrict__ offsets,
                                                  uint64_t *__restrict__ pairs) {
  const SetParams &lp = lv.sets[lset];
  const SetParams &rp = rv.sets[rset];
  const uint64_t li = (uint64_t)blockIdx.x * kQB + threadIdx.x;
  if (li >= lp.ent_cnt) return;
  const uint64_t j0 = lp.ent_off + li;  // storage position
  const uint32_t id0 = lv.ids[j0];
  double emn[D], emx[D], esz[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    emn[d] = lv.geom[(uint64_t)d * lv.n + j0];
    emx[d] = lv.geom[(uint64_t)(D + d) * lv.n + j0];
    esz[d] = emx[d] - emn[d];
  }
  uint64_t o = FILL ? offsets[li] : 0;
  uint32_t cnt = 0;
  uint32_t stack[StackCap<D>::value];
  int top = 0;
  if (rp.node_cnt != 0) stack[top++] = (uint32_t)rp.node_off;
  while (top > 0) {
    const uint32_t node = stack[--top];
    const double *nm = &rv.nmin[(uint64_t)node * D];
    const double *ns = &rv.nsize[(uint64_t)node * D];
    bool overlap = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double maxL = nm[d] + ns[d];
      const double maxR = emn[d] + esz[d];
      if (maxL <= emn[d] - tol || maxR <= nm[d] - tol) overlap = false;
    }
    if (!overlap) continue;
    const NodeInfo ni = rv.info[node];
    for (uint32_t j = ni.ebegin; j < ni.ebegin + ni.ecount; ++j) {
      const uint32_t id1 = rv.ids[j];
      if (SELF && id1 <= id0) continue;
      double fmn[D], fmx[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        fmn[d] = rv.geom[(uint64_t)d * rv.n + j];
        fmx[d] = rv.geom[(uint64_t)(D + d) * rv.n + j];
      }
      if (boxes_overlap_strict<D>(emn, emx, fmn, fmx, tol)) {
        if (FILL) pairs[o + cnt] = ((uint64_t)id0 << 32) | id1;
        ++cnt;
      }
    }
    for (uint32_t c = 0; c < ni.ccount; ++c) stack[top++] = ni.cbegin + c;
  }
  if (!FILL) counts[li] = cnt;
}

// ---------------------------------------------------------------------------
// query ordering: (set, Hilbert index (2-D / 3-D) or Morton code of the query's
// anchor point) so that the lanes of a wavefront walk the same nodes and read the same
// leaves.  The order decides speed only, never results (C3: the Hilbert order took the
// kNN launch from 29.8 to 29.3 ms against the Morton order).
// ---------------------------------------------------------------------------
// Hilbert index of a cell of a 2^b grid (Skilling, "Programming the Hilbert curve",
// AIP Conf. Proc. 707, 2004: axes to transposed index, then the transposed bits read
// level by level).
template <int D>
__device__ __forceinline__ uint64_t hilbert_index(uint32_t (&X)[D], uint32_t b) {
  const uint32_t top = 1u << (b - 1);
  for (uint32_t Q = top; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
#pragma unroll
  for (int i = 1; i < D; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = top; Q > 1; Q >>= 1)
    if (X[D - 1] & Q) t ^= Q - 1;
#pragma unroll
  for (int i = 0; i < D; ++i) X[i] ^= t;
  uint64_t h = 0;
  for (int j = (int)b - 1; j >= 0; --j)
#pragma unroll
    for (int i = 0; i < D; ++i) h = (h << 1) | ((X[i] >> j) & 1u);
  return h;
}

template <int D>
__global__ void __launch_bounds__(256) k_query_keys(const double *__restrict__ q, uint32_t stride,
                                                     const uint32_t *__restrict__ qset, uint64_t nq,
                                                     const TreeView *__restrict__ views,
                                                     const uint32_t *__restrict__ qview, uint32_t cbits,
                                                     uint32_t vshift, uint32_t drop, uint64_t *__restrict__ keys,
                                                     uint32_t *__restrict__ idx) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nq) return;
  const uint32_t s = qset ? qset[i] : 0u;
  const uint32_t v = qview ? qview[i] : 0u;
  const SetParams &sp = views[v].sets[s];
  const uint32_t M = sp.max_depth;
  const double top = (double)((1ull << M) - 1);
  uint32_t g[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double x = (q[i * stride + d] - sp.space_min[d]) * sp.raster[d];
    g[d] = (x > 0.0) ? (uint32_t)(x < top ? x : top) : 0u;
  }
  if constexpr (D <= 3) {
    // 2-D / 3-D: the Hilbert index of the leading L levels (no jumps between
    // neighbouring cells, so a wavefront's queries stay close), left-aligned
    const uint32_t L = M < 30u / D ? M : 30u / D;
    uint32_t X[D];
#pragma unroll
    for (int d = 0; d < D; ++d) X[d] = g[d] >> (M - L);
    const uint64_t h = L ? hilbert_index<D>(X, L) : 0ull;
    keys[i] = ((uint64_t)v << vshift) | ((uint64_t)s << cbits) | (h << (cbits - D * L));
  } else {
    // the code aligned to the widest tree of the launch, then its leading bits kept
    const uint32_t up = cbits + drop - D * M;
    keys[i] = ((uint64_t)v << vshift) | ((uint64_t)s << cbits) | ((morton_encode<D>(g) << up) >> drop);
  }
  idx[i] = (uint32_t)i;
}

// Dynamic LDS that caps a one-wavefront-block kernel at `waves` wavefronts per SIMD
// (4 SIMDs per CU share the CU's LDS): 0 = no cap.
static size_t occupancy_pad(Ctx *ctx, const void *kernel, uint32_t waves) {
  if (!waves) return 0;
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, std::pair<size_t, size_t>> cache;  // (lds per CU, static)
  std::pair<size_t, size_t> v;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({ctx->device, kernel});
    if (it == cache.end()) {
      hipDeviceProp_t prop;
      OTG_HIP_CHECK(hipGetDeviceProperties(&prop, ctx->device));
      hipFuncAttributes fa;
      OTG_HIP_CHECK(hipFuncGetAttributes(&fa, kernel));
      it = cache.emplace(std::make_pair(ctx->device, kernel),
                         std::make_pair((size_t)prop.maxSharedMemoryPerMultiProcessor, (size_t)fa.sharedSizeBytes))
               .first;
    }
    v = it->second;
  }
  const size_t per_block = v.first / (4 * (size_t)waves);
  return per_block > v.second ? per_block - v.second : 0;
}

// Blocks of a persistent launch: as many as are resident at once (never more than
// the tiles).
static uint32_t persistent_grid(Ctx *ctx, const void *kernel, int block, size_t shmem, uint32_t ntiles) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void *, size_t>, uint32_t> cache;
  uint32_t resident;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({ctx->device, kernel, shmem});
    if (it == cache.end()) {
      int per_cu = 0;
      OTG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, shmem));
      hipDeviceProp_t prop;
      OTG_HIP_CHECK(hipGetDeviceProperties(&prop, ctx->device));
      it = cache.emplace(std::make_tuple(ctx->device, kernel, shmem),
                         (uint32_t)std::max(1, per_cu) * (uint32_t)prop.multiProcessorCount).first;
    }
    resident = it->second;
  }
  return std::max<uint32_t>(1, std::min(resident, ntiles));
}

__global__ void k_fill_u32(uint32_t *__restrict__ a, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// The trees one launch serves: device TreeViews, each query's view index
// (null: a single tree) and the widest key layout among them.
struct ViewSet {
  const TreeView *views = nullptr;
  const uint32_t *qview = nullptr;
  uint32_t nviews = 1;
  uint32_t cbits = 0, sbits = 0;
  uint64_t max_nodes = 0, max_entities = 0;  // largest tree of the launch (bounds of ray walks)
  uint64_t max_set = 0;                        // largest set: local ids are below it
};

static uint32_t bit_width32(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }

constexpr uint64_t kOrderMinQueries = 16384;

template <int D>
static Buf<uint32_t> query_order(Ctx *ctx, const ViewSet &vs, const double *q, uint32_t stride, const uint32_t *qs,
                                 uint64_t nq, Arena *ar = nullptr) {
  // below this many queries the ordering (key pass + radix sort, about 20 launches)
  // costs more than the locality it buys: the queries run in input order
  if (nq < kOrderMinQueries || !ctx->tuning.order_queries) return Buf<uint32_t>();
  Buf<uint32_t> idx(ctx, ar, nq);
  // the leading 30 bits of the Morton code (10 levels in 3-D) give the wavefronts
  // their locality; the order decides speed only, never results
  const uint32_t drop = vs.cbits > 30 ? vs.cbits - 30 : 0;
  const uint32_t cbits = vs.cbits - drop;
  const uint32_t vbits = bit_width32(vs.nviews - 1);
  const uint32_t vshift = cbits + vs.sbits;
  if (vshift + vbits > 64) throw Error{OTG_ERR_UNSUPPORTED, "query ordering key exceeds 64 bits"};
  Buf<uint64_t> keys(ctx, ar, nq);
  launch(ctx, "query_keys", k_query_keys<D>, dim3(div_up(nq, 256)), dim3(256), 0, q, stride, qs, nq, vs.views,
         vs.qview, cbits, vshift, drop, keys.p, idx.p);
  radix_sort_pairs(ctx, keys.p, idx.p, nq, vshift + vbits, ar);
  return idx;
}

// ---------------------------------------------------------------------------
// CSR helpers
// ---------------------------------------------------------------------------
__global__ void k_offsets_u64(const uint32_t *__restrict__ scan, uint64_t n, uint32_t total, uint64_t *__restrict__ off) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) off[i] = scan[i];
  if (i == n) off[n] = total;
}

// One thread per (query, slot): consecutive threads read consecutive slots and
// write consecutive CSR positions, so both sides coalesce.
__global__ void k_compact_fixed(const uint32_t *__restrict__ src, uint32_t stride, const uint64_t *__restrict__ off,
                                uint64_t nq, uint32_t *__restrict__ dst) {
  const uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t qi = x / stride;
  const uint32_t t = (uint32_t)(x - qi * stride);
  if (qi >= nq) return;
  const uint64_t o = off[qi];
  // longer lists (kNN and ray bands) are written by their own pass
  if (t < off[qi + 1] - o) dst[o + t] = src[x];
}

__global__ void k_split_pairs(const uint64_t *__restrict__ keys, uint64_t np, uint32_t jbits, uint32_t *__restrict__ ids,
                              uint64_t *__restrict__ off) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < np) {
    uint64_t k = keys[i];
    ids[2 * i] = (uint32_t)(k >> jbits);
    ids[2 * i + 1] = (uint32_t)(k & ((1ull << jbits) - 1));
  }
  if (i <= np) off[i] = 2 * i;
}

__global__ void k_pack_pairs(uint64_t *__restrict__ pairs, uint64_t np, uint32_t jbits) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  uint64_t k = pairs[i];
  pairs[i] = ((k >> 32) << jbits) | (k & 0xFFFFFFFFull);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
Result::~Result() {
  if (ctx) {
    if (d_offsets) ctx->free(d_offsets);
    if (d_ids) ctx->free(d_ids);
  }
}

struct DevInput {
  Buf<double> own;
  Buf<uint32_t> own_sets;
  const double *q = nullptr;
  const uint32_t *sets = nullptr;
};

// Largest per-query set id of a device array (validation before any walk reads it).
__global__ void k_max_u32(const uint32_t *__restrict__ a, uint64_t n, uint32_t *__restrict__ out) {
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    m = a[i] > m ? a[i] : m;
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// Queries (and per-query set ids) in device memory: host arrays are copied.  Every
// set id must name a set of the tree (nsets): host ids are checked here, device ids
// by a max-reduction read back before any walk indexes the set table with them.
static void stage_inputs(Ctx *ctx, DevInput &in, const double *q, const uint32_t *sets, uint64_t count_doubles,
                         uint64_t nq, int mem, uint32_t nsets) {
  if (sets && nq) {
    uint32_t mx = 0;
    if (mem == OTG_MEM_DEVICE) {
      Buf<uint32_t> dm(ctx, 1);
      OTG_HIP_CHECK(hipMemsetAsync(dm.p, 0, 4, ctx->stream));
      launch(ctx, "check_sets", k_max_u32, dim3(std::min<uint64_t>(div_up(nq, 256), 1024)), dim3(256), 0, sets, nq,
             dm.p);
      OTG_HIP_CHECK(hipMemcpyAsync(&mx, dm.p, 4, hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
    } else {
      mx = *std::max_element(sets, sets + nq);
    }
    if (mx >= nsets) throw Error{OTG_ERR_INVALID_ARG, "a query's set id is not a set of the tree"};
  }
  if (mem == OTG_MEM_DEVICE) {
    in.q = q;
    in.sets = sets;
    return;
  }
  in.own = Buf<double>(ctx, count_doubles);
  if (count_doubles)
    OTG_HIP_CHECK(hipMemcpyAsync(in.own.p, q, count_doubles * 8, hipMemcpyHostToDevice, ctx->stream));
  in.q = in.own.p;
  if (sets) {
    in.own_sets = Buf<uint32_t>(ctx, nq);
    OTG_HIP_CHECK(hipMemcpyAsync(in.own_sets.p, sets, nq * 4, hipMemcpyHostToDevice, ctx->stream));
    in.sets = in.own_sets.p;
  }
}

// Result lists are indexed with 32-bit positions: a batch whose ids total 2^32 - 1
// or more (the scan reports it saturated) is rejected before anything is sized from it.
static void check_id_total(uint32_t total, Result *owned = nullptr) {
  if (total == kScanOverflow) {
    delete owned;
    throw Error{OTG_ERR_UNSUPPORTED, "a batch result of 2^32 - 1 or more ids (split the batch)"};
  }
}

// CSR offsets of a batch in which every query has k ids
__global__ void k_offsets_linear(uint64_t nq, uint32_t k, uint64_t *__restrict__ off) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i <= nq) off[i] = i * k;
}

// rows of the lists that had exactly k ids, back into the KCAP-slot lists
__global__ void k_direct_to_fixed(const uint32_t *__restrict__ direct, const uint32_t *__restrict__ cnt, uint64_t nq,
                                  uint32_t k, uint32_t kcap, uint32_t *__restrict__ fixed) {
  const uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t qi = x / k;
  if (qi >= nq || cnt[qi] != k) return;
  fixed[qi * kcap + (x - qi * k)] = direct[x];
}

// Fixed-stride per-query results -> CSR
static Result *fixed_to_csr(Ctx *ctx, const uint32_t *ids, const uint32_t *cnt, uint32_t stride, uint64_t nq) {
  Result *r = new Result();
  r->ctx = ctx;
  r->nq = nq;
  Buf<uint32_t> scan(ctx, nq + 1);
  Buf<uint32_t> total(ctx, 1);
  exclusive_scan_u32(ctx, cnt, scan.p, nq, total.p);
  uint32_t ht = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&ht, total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(ht, r);
  r->nids = ht;
  r->d_offsets = (uint64_t *)ctx->alloc((nq + 1) * 8);
  r->d_ids = (uint32_t *)ctx->alloc(std::max<uint64_t>(ht, 1) * 4);
  launch(ctx, "offsets_u64", k_offsets_u64, dim3(div_up(nq + 1, 256)), dim3(256), 0, (const uint32_t *)scan.p, nq, ht,
         r->d_offsets);
  if (stride)
    launch(ctx, "compact_fixed", k_compact_fixed, dim3(div_up(nq * stride, 256)), dim3(256), 0, ids, stride,
           (const uint64_t *)r->d_offsets, nq, r->d_ids);
  return r;
}

// footprint (optional, measurement): out[0..1] += node and entity visits of the minimal walk
__global__ void k_widen_u32(const uint32_t *__restrict__ in, uint64_t n, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    keys[i] = in[i];
    vals[i] = in[i];
  }
}

// Sorts a list of query-order positions appended by atomics, so a redo pass walks
// its queries in the (set, Morton) order again.
static void sort_positions(Ctx *ctx, uint32_t *list, uint32_t n, uint64_t nq) {
  if (n < 2) return;
  Buf<uint64_t> keys(ctx, n);
  launch(ctx, "redo_widen", k_widen_u32, dim3(div_up(n, 256)), dim3(256), 0, (const uint32_t *)list, (uint64_t)n,
         keys.p, list);
  radix_sort_pairs(ctx, keys.p, list, n, bit_width32((uint32_t)std::max<uint64_t>(nq - 1, 1)));
}

// Counts -> CSR offsets and compacted ids in the caller's buffers (the
// allocation-free path); nothing is written past the ids capacity.
static void csr_into(Ctx *ctx, Arena *ar, const uint32_t *ids, const uint32_t *cnt, uint32_t stride, uint64_t nq,
                     CsrOut &co) {
  Buf<uint32_t> scan(ctx, ar, nq + 1), total(ctx, ar, 1);
  exclusive_scan_u32(ctx, cnt, scan.p, nq, total.p, ar);
  uint32_t ht = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&ht, total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(ht);
  co.nids = ht;
  if (ht > co.capacity) throw Error{OTG_ERR_CAPACITY, "the result holds more ids than the output capacity"};
  launch(ctx, "offsets_u64", k_offsets_u64, dim3(div_up(nq + 1, 256)), dim3(256), 0, (const uint32_t *)scan.p, nq, ht,
         co.offsets);
  launch(ctx, "compact_fixed", k_compact_fixed, dim3(div_up(nq * stride, 256)), dim3(256), 0, ids, stride,
         (const uint64_t *)co.offsets, nq, co.ids);
}

// ar / co: the allocation-free path -- every temporary from the arena, the result
// into the caller's CSR (returns null), queries in input order.
template <int D, int KCAP, bool BOX>
static Result *knn_impl(Ctx *ctx, const ViewSet &vs, uint64_t nq, const double *q, const uint32_t *qs,
                        const uint32_t *qk, uint32_t k, double F0, double tol, uint64_t *footprint = nullptr,
                        Arena *ar = nullptr, CsrOut *co = nullptr) {
  Buf<uint32_t> order = ar ? Buf<uint32_t>() : query_order<D>(ctx, vs, q, D, qs, nq);
  Buf<uint32_t> ids(ctx, ar, nq * KCAP), cnt(ctx, ar, nq), ovf(ctx, ar, 2), band(ctx, ar, nq);
  Buf<double> Fq(ctx, ar, nq);
  OTG_HIP_CHECK(hipMemsetAsync(ovf.p, 0, 8, ctx->stream));  // band count, irregular list length
  // one k for the whole batch (the plain call): lists of exactly k ids go straight into
  // the result's rows
  Buf<uint32_t> direct = (!ar && !qk) ? Buf<uint32_t>(ctx, nq * k) : Buf<uint32_t>();
  const KnnOut kout{ids.p, direct.p, ovf.p + 1};
  bool waved = false;
  if constexpr (kKnnWave<D, KCAP, BOX>) {
    if (ctx->tuning.knn_wave && knn_wave_batch<D>(nq)) {
      launch(ctx, "knn", k_knn_wave<D, KCAP>, dim3(div_up(nq, kWaveQueries)), dim3(64 * kWaveQueries), 0, vs.views,
             vs.qview, q, qs, qk, (const uint32_t *)order.p, nq, k, F0, tol, kout, cnt.p, ovf.p, band.p, Fq.p);
      waved = true;
    }
  }
  if (!waved) {
    if (ctx->tuning.persistent) {
      const void *kp = (const void *)k_knn_persistent<D, KCAP, BOX>;
      const size_t pad = occupancy_pad(ctx, kp, ctx->tuning.waves_per_simd);
      Buf<uint32_t> tiles(ctx, ar, 1);
      launch(ctx, "fill_u32", k_fill_u32, dim3(1), dim3(64), 0, tiles.p, (uint64_t)1, 0u);
      launch(ctx, "knn", k_knn_persistent<D, KCAP, BOX>, dim3(persistent_grid(ctx, kp, kQB, pad, div_up(nq, kQB))),
             dim3(kQB), pad, vs.views, vs.qview, q, qs, qk, (const uint32_t *)order.p, nq, k, F0, tol, kout, cnt.p,
             ovf.p, band.p, Fq.p, tiles.p);
    } else {
      const size_t pad = occupancy_pad(ctx, (const void *)k_knn<D, KCAP, BOX>, ctx->tuning.waves_per_simd);
      launch(ctx, "knn", k_knn<D, KCAP, BOX>, dim3(div_up(nq, kQB)), dim3(kQB), pad, vs.views, vs.qview, q, qs, qk,
             (const uint32_t *)order.p, nq, k, F0, tol, kout, cnt.p, ovf.p, band.p, Fq.p);
    }
  }
  uint32_t hovf[2] = {0, 0};
  OTG_HIP_CHECK(hipMemcpyAsync(hovf, ovf.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  if (footprint) {
    Buf<unsigned long long> acc(ctx, 2);
    OTG_HIP_CHECK(hipMemsetAsync(acc.p, 0, 16, ctx->stream));
    launch(ctx, "knn_footprint", k_knn_footprint<D>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views, vs.qview, q, qs,
           (const uint32_t *)order.p, nq, (const double *)Fq.p, acc.p);
    unsigned long long h[2] = {0, 0};
    OTG_HIP_CHECK(hipMemcpyAsync(h, acc.p, 16, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    footprint[0] += h[0];
    footprint[1] += h[1];
  }
  const uint32_t nb = hovf[0];
  if (direct.p) {
    if (nb == 0 && hovf[1] == 0) {  // every query has exactly k ids: the rows are the result
      Result *r = new Result();
      r->ctx = ctx;
      r->nq = nq;
      r->nids = nq * k;
      check_id_total(r->nids > 0xFFFFFFFEull ? kScanOverflow : (uint32_t)r->nids, r);
      r->d_offsets = (uint64_t *)ctx->alloc((nq + 1) * 8);
      launch(ctx, "offsets_linear", k_offsets_linear, dim3(div_up(nq + 1, 256)), dim3(256), 0, nq, k, r->d_offsets);
      r->d_ids = direct.release();
      return r;
    }
    // lists of other lengths: the rows join the fixed lists, then the general path
    launch(ctx, "direct_to_fixed", k_direct_to_fixed, dim3(div_up(nq * k, 256)), dim3(256), 0,
           (const uint32_t *)direct.p, (const uint32_t *)cnt.p, nq, k, (uint32_t)KCAP, ids.p);
  }
  if (nb == 0) {
    if (co) {
      csr_into(ctx, ar, ids.p, cnt.p, KCAP, nq, *co);
      return nullptr;
    }
    return fixed_to_csr(ctx, ids.p, cnt.p, KCAP, nq);
  }
  // band queries: count, fill, sort, then splice into the CSR
  Buf<uint32_t> bcnt(ctx, ar, nb), boff(ctx, ar, nb + 1), btot(ctx, ar, 1);
  launch(ctx, "knn_band_count", k_knn_band<D, BOX, false>, dim3(div_up(nb, kQB)), dim3(kQB), 0, vs.views, vs.qview, q,
         qs, (const uint32_t *)band.p, nb, (const double *)Fq.p, tol, bcnt.p, (const uint32_t *)nullptr,
         (double *)nullptr, (uint32_t *)nullptr);
  exclusive_scan_u32(ctx, bcnt.p, boff.p, nb, btot.p, ar);
  uint32_t hb = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&hb, btot.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(hb);
  OTG_HIP_CHECK(hipMemcpyAsync(boff.p + nb, &hb, 4, hipMemcpyHostToDevice, ctx->stream));
  Buf<double> bd(ctx, ar, std::max<uint32_t>(hb, 1));
  Buf<uint32_t> bi(ctx, ar, std::max<uint32_t>(hb, 1));
  launch(ctx, "knn_band_fill", k_knn_band<D, BOX, true>, dim3(div_up(nb, kQB)), dim3(kQB), 0, vs.views, vs.qview, q,
         qs, (const uint32_t *)band.p, nb, (const double *)Fq.p, tol, (uint32_t *)nullptr, (const uint32_t *)boff.p,
         bd.p, bi.p);
  launch(ctx, "knn_band_sort", k_sort_band<false, uint32_t>, dim3(div_up(nb, kQB)), dim3(kQB), 0, (const uint32_t *)boff.p, (uint64_t)nb,
         bd.p, bi.p);
  launch(ctx, "knn_band_counts", k_band_counts, dim3(div_up(nb, 256)), dim3(256), 0, (const uint32_t *)band.p, nb,
         (const uint32_t *)bcnt.p, cnt.p);
  Result *r = nullptr;
  uint64_t *offsets;
  uint32_t *out_ids;
  if (co) {
    csr_into(ctx, ar, ids.p, cnt.p, KCAP, nq, *co);
    offsets = co->offsets;
    out_ids = co->ids;
  } else {
    r = fixed_to_csr(ctx, ids.p, cnt.p, KCAP, nq);
    offsets = r->d_offsets;
    out_ids = r->d_ids;
  }
  launch(ctx, "knn_band_copy", k_band_copy, dim3(div_up(nb, 256)), dim3(256), 0, (const uint32_t *)band.p, nb,
         (const uint32_t *)boff.p, (const uint32_t *)bi.p, (const uint64_t *)offsets, out_ids);
  ctx->sync();  // the staged count above lives on this stack frame
  return r;
}

// k > 64: the radius pass in chunks (k doubles of heap per query, at most 256 MiB at
// a time), then the band pass over every query with that radius gives the lists.
template <int D, bool BOX>
static Result *knn_large_k(Ctx *ctx, const ViewSet &vs, uint64_t nq, const double *q, const uint32_t *qs, uint32_t k,
                           double F0, double tol) {
  Buf<double> Fq(ctx, nq);
  const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(nq, (256ull << 20) / (8ull * k)));
  {
    Buf<double> heap(ctx, chunk * k);
    for (uint64_t q0 = 0; q0 < nq; q0 += chunk) {
      const uint64_t m = std::min(chunk, nq - q0);
      launch(ctx, "knn_radius", k_knn_radius<D, BOX>, dim3(div_up(m, kQB)), dim3(kQB), 0, vs.views, vs.qview, q, qs,
             q0, m, k, F0, tol, heap.p, Fq.p);
    }
  }
  if (nq > 0xFFFFFFFFull) throw Error{OTG_ERR_UNSUPPORTED, "more than 2^32 - 1 queries with k > 64"};
  const uint32_t nb = (uint32_t)nq;
  Buf<uint32_t> all(ctx, nb), bcnt(ctx, nb), boff(ctx, nb + 1), btot(ctx, 1);
  launch(ctx, "iota_u32", k_iota_u32, dim3(div_up(nb, 256)), dim3(256), 0, all.p, nb);
  launch(ctx, "knn_band_count", k_knn_band<D, BOX, false>, dim3(div_up(nb, kQB)), dim3(kQB), 0, vs.views, vs.qview, q,
         qs, (const uint32_t *)all.p, nb, (const double *)Fq.p, tol, bcnt.p, (const uint32_t *)nullptr,
         (double *)nullptr, (uint32_t *)nullptr);
  exclusive_scan_u32(ctx, bcnt.p, boff.p, nb, btot.p);
  uint32_t hb = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&hb, btot.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(hb);
  OTG_HIP_CHECK(hipMemcpyAsync(boff.p + nb, &hb, 4, hipMemcpyHostToDevice, ctx->stream));
  Buf<double> bd(ctx, std::max<uint32_t>(hb, 1));
  Result *r = new Result();
  r->ctx = ctx;
  r->nq = nq;
  r->nids = hb;
  r->d_ids = (uint32_t *)ctx->alloc(std::max<uint64_t>(hb, 1) * 4);
  r->d_offsets = (uint64_t *)ctx->alloc((nq + 1) * 8);
  launch(ctx, "knn_band_fill", k_knn_band<D, BOX, true>, dim3(div_up(nb, kQB)), dim3(kQB), 0, vs.views, vs.qview, q,
         qs, (const uint32_t *)all.p, nb, (const double *)Fq.p, tol, (uint32_t *)nullptr, (const uint32_t *)boff.p,
         bd.p, r->d_ids);
  launch(ctx, "knn_band_sort", k_sort_band<false, uint32_t>, dim3(div_up(nb, kQB)), dim3(kQB), 0,
         (const uint32_t *)boff.p, (uint64_t)nb, bd.p, r->d_ids);
  launch(ctx, "offsets_u64", k_offsets_u64, dim3(div_up(nq + 1, 256)), dim3(256), 0, (const uint32_t *)boff.p, nq, hb,
         r->d_offsets);
  ctx->sync();  // the staged total above lives on this stack frame
  return r;
}

static Result *knn_large_k_dispatch(Ctx *ctx, int dim, bool box, const ViewSet &vs, uint64_t nq, const double *q,
                                    const uint32_t *qs, uint32_t k, double F0, double tol) {
#define OTG_KNNL(DD)                                                                        \
  case DD:                                                                                  \
    return box ? knn_large_k<DD, true>(ctx, vs, nq, q, qs, k, F0, tol)                      \
               : knn_large_k<DD, false>(ctx, vs, nq, q, qs, k, F0, tol);
  switch (dim) {
    OTG_KNNL(2)
    OTG_KNNL(3)
    OTG_KNNL(4)
    OTG_KNNL(5)
    OTG_KNNL(6)
  }
#undef OTG_KNNL
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

// kmax = the largest k of the launch (per-query k in qk, or k for all)
static Result *knn_dispatch(Ctx *ctx, int dim, bool box, const ViewSet &vs, uint64_t nq, const double *q,
                            const uint32_t *qs, const uint32_t *qk, uint32_t k, uint32_t kmax, double F0, double tol,
                            uint64_t *footprint = nullptr, Arena *ar = nullptr, CsrOut *co = nullptr) {
// list capacity k + 1 for the largest k of a class: one spare slot holds an entry of the
// tolerance band above the k-th distance; a longer band is redone by the band pass
#define OTG_KNN_K(DD, BB)                                                                                  \
  if (kmax <= 8) return knn_impl<DD, 9, BB>(ctx, vs, nq, q, qs, qk, k, F0, tol, footprint, ar, co);           \
  if (kmax <= 16) return knn_impl<DD, 17, BB>(ctx, vs, nq, q, qs, qk, k, F0, tol, footprint, ar, co);         \
  if (kmax <= 32) return knn_impl<DD, 33, BB>(ctx, vs, nq, q, qs, qk, k, F0, tol, footprint, ar, co);         \
  return knn_impl<DD, 65, BB>(ctx, vs, nq, q, qs, qk, k, F0, tol, footprint, ar, co);
#define OTG_KNN(DD)          \
  case DD:                   \
    if (box) {               \
      OTG_KNN_K(DD, true)    \
    } else {                 \
      OTG_KNN_K(DD, false)   \
    }
  switch (dim) {
    OTG_KNN(2)
    OTG_KNN(3)
    OTG_KNN(4)
    OTG_KNN(5)
    OTG_KNN(6)
  }
#undef OTG_KNN
#undef OTG_KNN_K
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

static Result *empty_result(Ctx *ctx, uint64_t nq) {
  Buf<uint32_t> z(ctx, std::max<uint64_t>(nq, 1));
  OTG_HIP_CHECK(hipMemsetAsync(z.p, 0, std::max<uint64_t>(nq, 1) * 4, ctx->stream));
  return fixed_to_csr(ctx, z.p, z.p, 0, nq);
}

static ViewSet single_view(const Tree *t) {
  ViewSet vs;
  vs.views = t->d_view;
  vs.cbits = t->layout.cbits;
  vs.sbits = t->layout.sbits;
  vs.max_nodes = t->node_count;
  vs.max_entities = t->n;
  for (const SetParams &sp : t->sets) vs.max_set = std::max<uint64_t>(vs.max_set, sp.ent_cnt);
  return vs;
}

// kNN answers come from the finer index when there is one (the kNN set depends on distances only).
static const Tree *knn_index(const Tree *t) { return t->knn_tree ? t->knn_tree : t; }
// Point range answers too: each entity is tested, or included with a node box
// that the range contains (ot_query.h:543-580), whichever tree is walked.
static const Tree *range_index(const Tree *t) {
  return (t->knn_tree && t->kind == OTG_KIND_POINT) ? t->knn_tree : t;
}

static double knn_F0(double max_within, double tol) {
  return (max_within == DBL_MAX) ? DBL_MAX : fmax(tol, max_within * (1.0 + tol));
}

Result *query_knn(const Tree *t, uint64_t nq, const double *pts, const uint32_t *sets, int mem, uint32_t k,
                  double max_within, double tol, uint64_t *footprint) {
  if (k > 64 && footprint) throw Error{OTG_ERR_UNSUPPORTED, "footprint of k > 64"};
  Ctx *ctx = run_ctx(t);
  t = knn_index(t);
  DevInput in;
  stage_inputs(ctx, in, pts, sets, nq * t->dim, nq, mem, (uint32_t)t->sets.size());
  if (nq == 0 || t->n == 0 || k == 0) return empty_result(ctx, nq);
  if (k > 64)
    return knn_large_k_dispatch(ctx, t->dim, t->kind == OTG_KIND_BOX, single_view(t), nq, in.q, in.sets, k,
                                knn_F0(max_within, tol), tol);
  return knn_dispatch(ctx, t->dim, t->kind == OTG_KIND_BOX, single_view(t), nq, in.q, in.sets, nullptr, k, k,
                      knn_F0(max_within, tol), tol, footprint);
}

void query_knn_noalloc(const Tree *t, uint64_t nq, const double *pts, const uint32_t *sets, uint32_t k,
                       double max_within, double tol, Arena &ar, CsrOut &co) {
  if (k > 64) throw Error{OTG_ERR_UNSUPPORTED, "k > 64"};
  Ctx *ctx = run_ctx(t);
  const Tree *x = knn_index(t);
  if (sets && nq) {
    Buf<uint32_t> dm(&ar, 1);
    OTG_HIP_CHECK(hipMemsetAsync(dm.p, 0, 4, ctx->stream));
    launch(ctx, "check_sets", k_max_u32, dim3(std::min<uint64_t>(div_up(nq, 256), 1024)), dim3(256), 0, sets, nq,
           dm.p);
    uint32_t mx = 0;
    OTG_HIP_CHECK(hipMemcpyAsync(&mx, dm.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (mx >= x->sets.size()) throw Error{OTG_ERR_INVALID_ARG, "a query's set id is not a set of the tree"};
  }
  co.nids = 0;
  if (nq == 0 || x->n == 0 || k == 0) {
    OTG_HIP_CHECK(hipMemsetAsync(co.offsets, 0, (nq + 1) * 8, ctx->stream));
  } else {
    knn_dispatch(ctx, x->dim, x->kind == OTG_KIND_BOX, single_view(x), nq, pts, sets, nullptr, k, k,
                 knn_F0(max_within, tol), tol, nullptr, &ar, &co);
  }
  ctx->sync();
}

// One-lane walks stop counting after this many node visits; such queries, and those
// with more than kRangeInline ids, are walked by a wavefront each (2-D / 3-D).
constexpr uint32_t kRangeVisitCap = 32;

// Positions (in walk order) of the flagged queries, as query indices.
static uint32_t flagged_list(Ctx *ctx, const uint32_t *flags, const uint32_t *order, uint64_t nq, Buf<uint32_t> &list) {
  Buf<uint32_t> pos(ctx, nq + 1), total(ctx, 1);
  exclusive_scan_u32(ctx, flags, pos.p, nq, total.p);
  uint32_t n = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&n, total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  list = Buf<uint32_t>(ctx, std::max<uint32_t>(n, 1));
  launch(ctx, "range_heavy_list", k_range_heavy_list, dim3(div_up(nq, 256)), dim3(256), 0, flags,
         (const uint32_t *)pos.p, nq, order, list.p);
  return n;
}

// Sorts the ids of the listed queries of a CSR result (qlist null: queries 0..n-1)
// ascending: one radix sort of (rank in the list, id) keys over all their ids.
// Bits of the local ids of a launch's trees (every id is below the largest set's size).
static uint32_t id_bits(const ViewSet &vs) {
  return std::max(1u, bit_width32((uint32_t)std::min<uint64_t>(std::max<uint64_t>(vs.max_set, 1) - 1, 0xFFFFFFFFull)));
}

static void sort_segments(Ctx *ctx, Result *r, const uint32_t *qlist, uint32_t n, uint32_t idbits) {
  if (n == 0) return;
  Buf<uint32_t> all;
  if (!qlist) {
    all = Buf<uint32_t>(ctx, n);
    launch(ctx, "iota_u32", k_iota_u32, dim3(div_up(n, 256)), dim3(256), 0, all.p, n);
    qlist = all.p;
  }
  Buf<uint32_t> hc(ctx, n), hpos(ctx, n + 1), htot(ctx, 1);
  launch(ctx, "range_heavy_counts", k_range_heavy_counts, dim3(div_up(n, 256)), dim3(256), 0, qlist, n,
         (const uint64_t *)r->d_offsets, hc.p, (const uint32_t *)nullptr);
  exclusive_scan_u32(ctx, hc.p, hpos.p, n, htot.p);
  uint32_t nh = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&nh, htot.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(nh);
  if (nh < 2) return;
  Buf<uint64_t> keys(ctx, nh);
  Buf<uint32_t> vals(ctx, nh);
  launch(ctx, "range_sort_keys", k_range_sort_keys, dim3(n), dim3(256), 0, qlist, n, (const uint64_t *)r->d_offsets,
         (const uint32_t *)hpos.p, (const uint32_t *)r->d_ids, keys.p, (const uint32_t *)nullptr, (const uint32_t *)nullptr,
         idbits);
  radix_sort_pairs(ctx, keys.p, vals.p, nh, idbits + std::max(1u, bit_width32(n - 1)));
  launch(ctx, "range_sort_back", k_range_sort_back, dim3(n), dim3(256), 0, qlist, n, (const uint64_t *)r->d_offsets,
         (const uint32_t *)hpos.p, (const uint64_t *)keys.p, r->d_ids, (const uint32_t *)nullptr, (const uint32_t *)nullptr,
         idbits);
}

template <int D, bool BOX>
static Result *range_impl(Ctx *ctx, const ViewSet &vs, uint64_t nq, const double *q, const uint32_t *qs, int mode,
                          double tol) {
  constexpr bool kWave = D <= 3;
  Buf<uint32_t> order = query_order<D>(ctx, vs, q, 2 * D, qs, nq);
  const uint32_t *ord = order.p;
  Buf<uint32_t> cnt(ctx, std::max<uint64_t>(nq, 1)), ncopy(ctx, 1);
  Buf<uint32_t> inl(ctx, std::max<uint64_t>(nq, 1) * kRangeInline);
  OTG_HIP_CHECK(hipMemsetAsync(ncopy.p, 0, 4, ctx->stream));
  // 1. one-lane count walks (ids of small results kept inline), capped for the wavefront path
  launch(ctx, "range_count", k_range<D, BOX, false>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views, vs.qview, q, qs,
         ord, nq, mode, tol, cnt.p, (const uint64_t *)nullptr, (uint32_t *)nullptr, ncopy.p, (RangeCopy *)nullptr,
         inl.p, kWave ? kRangeVisitCap : 0u);
  Buf<uint32_t> capped(ctx, std::max<uint64_t>(nq, 1));
  // the capped queries' lists kept by their count walk (bounded slots per query; a
  // query that outgrows them is filled by the fill walk instead)
  Buf<uint32_t> keep_n, keep_ok, keep_xq, clist;
  Buf<RangeRec> keep_recs;
  RangeKeep keep{};
  uint32_t ncap = 0;
  if constexpr (kWave) {
    // 2. the capped ones: counted by a wavefront each, their ids kept
    launch(ctx, "range_capped", k_range_capped, dim3(div_up(nq, 256)), dim3(256), 0, ord, nq, (const uint32_t *)cnt.p,
           capped.p);
    ncap = flagged_list(ctx, capped.p, ord, nq, clist);
    const uint32_t want = ctx->tuning.range_records;
    if (ncap && (uint64_t)ncap * (want ? want : kKeepMin) <= kKeepBudget) {
      const uint32_t kcap = want ? want : (uint32_t)std::min<uint64_t>(kKeepMax, kKeepBudget / ncap);
      const uint32_t xcap = (uint32_t)std::min<uint64_t>(kKeepOverflow, 1024ull * kcap);
      keep_recs = Buf<RangeRec>(ctx, (uint64_t)ncap * kcap + xcap);
      keep_n = Buf<uint32_t>(ctx, ncap + 1);
      keep_ok = Buf<uint32_t>(ctx, std::max<uint64_t>(nq, 1));
      keep_xq = Buf<uint32_t>(ctx, xcap);
      OTG_HIP_CHECK(hipMemsetAsync(keep_n.p + ncap, 0, 4, ctx->stream));
      keep = RangeKeep{keep_recs.p, keep_n.p, keep_ok.p, keep_recs.p + (uint64_t)ncap * kcap,
                       keep_xq.p,   keep_n.p + ncap, kcap, xcap, ncap};
    }
    launch(ctx, "range_wave_count", k_range_wave<D, BOX, false>, dim3(ncap), dim3(64), 0, vs.views, vs.qview, q, qs,
           (const uint32_t *)clist.p, mode, tol, cnt.p, (const uint64_t *)nullptr, (uint32_t *)nullptr, ncopy.p,
           (RangeCopy *)nullptr, (const uint32_t *)nullptr, (const uint32_t *)nullptr, keep);
  } else {
    OTG_HIP_CHECK(hipMemsetAsync(capped.p, 0, std::max<uint64_t>(nq, 1) * 4, ctx->stream));
  }
  // 3. offsets
  Result *r = new Result();
  r->ctx = ctx;
  r->nq = nq;
  r->capped = ncap;
  Buf<uint32_t> scan(ctx, nq + 1), total(ctx, 1);
  exclusive_scan_u32(ctx, cnt.p, scan.p, nq, total.p);
  uint32_t ht = 0, hcopy = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&ht, total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  OTG_HIP_CHECK(hipMemcpyAsync(&hcopy, ncopy.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(ht, r);
  r->nids = ht;
  r->d_offsets = (uint64_t *)ctx->alloc((nq + 1) * 8);
  r->d_ids = (uint32_t *)ctx->alloc(std::max<uint64_t>(ht, 1) * 4);
  // hcopy counts the deferred runs of every count walk (a capped query's twice): a bound
  Buf<RangeCopy> copies(ctx, std::max<uint32_t>(hcopy, 1));
  OTG_HIP_CHECK(hipMemsetAsync(ncopy.p, 0, 4, ctx->stream));
  launch(ctx, "offsets_u64", k_offsets_u64, dim3(div_up(nq + 1, 256)), dim3(256), 0, (const uint32_t *)scan.p, nq, ht,
         r->d_offsets);
  // 4. small complete results from the inline rows; the kept ones from their slots;
  // the rest are filled by a second walk
  const bool collected = keep.recs != nullptr;
  Buf<uint32_t> heavy(ctx, std::max<uint64_t>(nq, 1)), walk;
  if (collected) walk = Buf<uint32_t>(ctx, std::max<uint64_t>(nq, 1));
  launch(ctx, "range_inline", k_range_inline, dim3(div_up(nq, 256)), dim3(256), 0, ord, nq,
         (const uint32_t *)cnt.p, (const uint32_t *)capped.p, (const uint64_t *)r->d_offsets, (const uint32_t *)inl.p,
         r->d_ids, heavy.p, (const uint32_t *)nullptr, walk.p, (const uint32_t *)keep.ok);
  Buf<uint32_t> hlist, wlist;
  const uint32_t nheavy = flagged_list(ctx, heavy.p, ord, nq, hlist);
  const uint32_t nwalk = collected ? flagged_list(ctx, walk.p, ord, nq, wlist) : nheavy;
  if (collected)
    launch(ctx, "range_kept", k_range_kept<D, BOX>, dim3(ncap + 256), dim3(256), 0, vs.views, vs.qview, q,
           (const uint32_t *)clist.p, ncap, mode, tol, keep, (const uint64_t *)r->d_offsets, r->d_ids,
           (const uint32_t *)nullptr, (const uint32_t *)nullptr);
  if constexpr (kWave)
    launch(ctx, "range_fill", k_range_wave<D, BOX, true>, dim3(nwalk), dim3(64), 0, vs.views, vs.qview, q, qs,
           (const uint32_t *)(collected ? wlist.p : hlist.p), mode, tol, (uint32_t *)nullptr,
           (const uint64_t *)r->d_offsets, r->d_ids, ncopy.p, copies.p, (const uint32_t *)nullptr,
           (const uint32_t *)nullptr, RangeKeep{});
  else
    launch(ctx, "range_fill", k_range<D, BOX, true>, dim3(div_up(nheavy, kQB)), dim3(kQB), 0, vs.views, vs.qview, q,
           qs, (const uint32_t *)hlist.p, (uint64_t)nheavy, mode, tol, (uint32_t *)nullptr,
           (const uint64_t *)r->d_offsets, r->d_ids, ncopy.p, copies.p, (uint32_t *)nullptr, 0u);
  // 5. long contained runs, one wavefront each
  OTG_HIP_CHECK(hipMemcpyAsync(&hcopy, ncopy.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  launch(ctx, "range_copy", k_range_copy, dim3(div_up(hcopy, 4)), dim3(256), 0, (const RangeCopy *)copies.p, hcopy,
         r->d_ids, (const uint32_t *)nullptr, (const uint32_t *)nullptr);
  // 6. the larger results sorted by id
  sort_segments(ctx, r, hlist.p, nheavy, id_bits(vs));
  return r;
}

static Result *range_dispatch(Ctx *ctx, int dim, bool box, const ViewSet &vs, uint64_t nq, const double *q,
                              const uint32_t *qs, int mode, double tol) {
#define OTG_RANGE(DD)                                                  \
  case DD:                                                             \
    return box ? range_impl<DD, true>(ctx, vs, nq, q, qs, mode, tol)   \
               : range_impl<DD, false>(ctx, vs, nq, q, qs, mode, tol);
  switch (dim) {
    OTG_RANGE(2)
    OTG_RANGE(3)
    OTG_RANGE(4)
    OTG_RANGE(5)
    OTG_RANGE(6)
  }
#undef OTG_RANGE
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

Result *query_range(const Tree *t, uint64_t nq, const double *ranges, const uint32_t *sets, int mem, int mode,
                    double tol) {
  Ctx *ctx = run_ctx(t);
  t = range_index(t);
  DevInput in;
  stage_inputs(ctx, in, ranges, sets, nq * 2 * t->dim, nq, mem, (uint32_t)t->sets.size());
  return range_dispatch(ctx, t->dim, t->kind == OTG_KIND_BOX, single_view(t), nq, in.q, in.sets, mode, tol);
}

// Count pass, CSR offsets, fill pass into the result: the shape shared by pick
// and ray-all.  count(counts) and fill(offsets, ids, dist) launch the walks.
template <typename Count, typename Fill>
static Result *count_fill_csr(Ctx *ctx, uint64_t nq, Count count, Fill fill) {
  Buf<uint32_t> cnt(ctx, std::max<uint64_t>(nq, 1));
  count(cnt.p);
  Result *r = new Result();
  r->ctx = ctx;
  r->nq = nq;
  Buf<uint32_t> scan(ctx, nq + 1), total(ctx, 1);
  exclusive_scan_u32(ctx, cnt.p, scan.p, nq, total.p);
  uint32_t ht = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&ht, total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(ht, r);
  r->nids = ht;
  r->d_offsets = (uint64_t *)ctx->alloc((nq + 1) * 8);
  r->d_ids = (uint32_t *)ctx->alloc(std::max<uint64_t>(ht, 1) * 4);
  launch(ctx, "offsets_u64", k_offsets_u64, dim3(div_up(nq + 1, 256)), dim3(256), 0, (const uint32_t *)scan.p, nq, ht,
         r->d_offsets);
  fill(r, ht);
  return r;
}

constexpr uint64_t kPlaneWaveQueries = 1u << 16;
constexpr uint64_t kPlaneWideQueries = 4096;

template <int D, bool BOX, int MODE>
static Result *planes_impl(Ctx *ctx, const ViewSet &vs, uint64_t nq, const double *q, uint32_t nplanes,
                           const uint32_t *qs, double tol) {
  // few queries, each a slab of the tree: one wavefront per query (2-D / 3-D)
  const bool wave = D <= 3 && nq < kPlaneWaveQueries;
  // up to a few thousand queries: four wavefronts per query so the chip fills
  const bool wide = nq <= kPlaneWideQueries;
  Buf<uint32_t> ncopy(ctx, 1);
  OTG_HIP_CHECK(hipMemsetAsync(ncopy.p, 0, 4, ctx->stream));
  Buf<RangeCopy> copies;
  uint32_t hcopy = 0;
  Result *r = count_fill_csr(
      ctx, nq,
      [&](uint32_t *cnt) {
        if constexpr (D <= 3) {
          if (wave) {
            if (wide)
              launch(ctx, "planes_count", k_planes_wave<D, BOX, MODE, false, 256>, dim3((uint32_t)nq), dim3(256), 0,
                     vs.views, vs.qview, q, nplanes, qs, nq, tol, cnt, (const uint64_t *)nullptr, (uint32_t *)nullptr,
                     ncopy.p, (RangeCopy *)nullptr);
            else
              launch(ctx, "planes_count", k_planes_wave<D, BOX, MODE, false>, dim3((uint32_t)nq), dim3(64), 0,
                     vs.views, vs.qview, q, nplanes, qs, nq, tol, cnt, (const uint64_t *)nullptr, (uint32_t *)nullptr,
                     ncopy.p, (RangeCopy *)nullptr);
            OTG_HIP_CHECK(hipMemcpyAsync(&hcopy, ncopy.p, 4, hipMemcpyDeviceToHost, ctx->stream));
            return;
          }
        }
        launch(ctx, "planes_count", k_planes<D, BOX, MODE, false>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views,
                 vs.qview, q, nplanes, qs, nq, tol, cnt, (const uint64_t *)nullptr, (uint32_t *)nullptr, ncopy.p,
                 (RangeCopy *)nullptr);
        OTG_HIP_CHECK(hipMemcpyAsync(&hcopy, ncopy.p, 4, hipMemcpyDeviceToHost, ctx->stream));
      },
      [&](Result *res, uint32_t) {
        copies = Buf<RangeCopy>(ctx, std::max<uint32_t>(hcopy, 1));
        OTG_HIP_CHECK(hipMemsetAsync(ncopy.p, 0, 4, ctx->stream));
        if constexpr (D <= 3) {
          if (wave) {
            if (wide)
              launch(ctx, "planes_fill", k_planes_wave<D, BOX, MODE, true, 256>, dim3((uint32_t)nq), dim3(256), 0,
                     vs.views, vs.qview, q, nplanes, qs, nq, tol, (uint32_t *)nullptr, (const uint64_t *)res->d_offsets,
                     res->d_ids, ncopy.p, copies.p);
            else
              launch(ctx, "planes_fill", k_planes_wave<D, BOX, MODE, true>, dim3((uint32_t)nq), dim3(64), 0, vs.views,
                     vs.qview, q, nplanes, qs, nq, tol, (uint32_t *)nullptr, (const uint64_t *)res->d_offsets,
                     res->d_ids, ncopy.p, copies.p);
            return;
          }
        }
        launch(ctx, "planes_fill", k_planes<D, BOX, MODE, true>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views,
                 vs.qview, q, nplanes, qs, nq, tol, (uint32_t *)nullptr, (const uint64_t *)res->d_offsets, res->d_ids,
                 ncopy.p, copies.p);
      });
  launch(ctx, "planes_copy", k_range_copy, dim3(div_up(hcopy, 4)), dim3(256), 0, (const RangeCopy *)copies.p, hcopy,
         r->d_ids, (const uint32_t *)nullptr, (const uint32_t *)nullptr);
  // the wavefront walks append in no fixed order: every list is returned sorted by id
  sort_segments(ctx, r, nullptr, (uint32_t)nq, id_bits(vs));
  return r;
}

Result *query_planes(const Tree *t, uint64_t nq, uint32_t nplanes, const double *planes, const uint32_t *sets, int mem,
                     int mode, double tol) {
  if (mode < 0 || mode > 2) throw Error{OTG_ERR_INVALID_ARG, "plane query mode"};
  if (mode != 2 && nplanes != 1) throw Error{OTG_ERR_INVALID_ARG, "plane search takes one plane per query"};
  // the reference tree: a node on the positive side contributes its whole subtree
  Ctx *ctx = run_ctx(t);
  DevInput in;
  stage_inputs(ctx, in, planes, sets, nq * nplanes * (t->dim + 1), nq, mem, (uint32_t)t->sets.size());
  const bool box = t->kind == OTG_KIND_BOX;
  const ViewSet vs = single_view(t);
#define OTG_PLANES_M(DD, BB)                                                                        \
  switch (mode) {                                                                                   \
    case 0: return planes_impl<DD, BB, 0>(ctx, vs, nq, in.q, nplanes, in.sets, tol);                \
    case 1: return planes_impl<DD, BB, 1>(ctx, vs, nq, in.q, nplanes, in.sets, tol);                \
    default: return planes_impl<DD, BB, 2>(ctx, vs, nq, in.q, nplanes, in.sets, tol);               \
  }
#define OTG_PLANES(DD)       \
  case DD:                   \
    if (box) {               \
      OTG_PLANES_M(DD, true) \
    } else {                 \
      OTG_PLANES_M(DD, false)\
    }
  switch (t->dim) {
    OTG_PLANES(2)
    OTG_PLANES(3)
    OTG_PLANES(4)
    OTG_PLANES(5)
    OTG_PLANES(6)
  }
#undef OTG_PLANES
#undef OTG_PLANES_M
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

template <int D, bool BOX>
static Result *pick_impl(Ctx *ctx, const ViewSet &vs, uint64_t nq, const double *q, const uint32_t *qs, double tol) {
  Buf<uint32_t> order = query_order<D>(ctx, vs, q, D, qs, nq);
  const uint32_t *ord = order.p;
  return count_fill_csr(
      ctx, nq,
      [&](uint32_t *cnt) {
        launch(ctx, "pick_count", k_pick<D, BOX, false>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views, vs.qview, q,
               qs, ord, nq, tol, cnt, (const uint64_t *)nullptr, (uint32_t *)nullptr);
      },
      [&](Result *r, uint32_t) {
        launch(ctx, "pick_fill", k_pick<D, BOX, true>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views, vs.qview, q, qs,
               ord, nq, tol, (uint32_t *)nullptr, (const uint64_t *)r->d_offsets, r->d_ids);
      });
}

Result *query_pick(const Tree *t, uint64_t nq, const double *picks, const uint32_t *sets, int mem, double tol) {
  // the reference tree: its node boxes decide the walk (ot_query.h:433-437)
  Ctx *ctx = run_ctx(t);
  DevInput in;
  stage_inputs(ctx, in, picks, sets, nq * t->dim, nq, mem, (uint32_t)t->sets.size());
  const bool box = t->kind == OTG_KIND_BOX;
  const ViewSet vs = single_view(t);
#define OTG_PICK(DD) \
  case DD: return box ? pick_impl<DD, true>(ctx, vs, nq, in.q, in.sets, tol) : pick_impl<DD, false>(ctx, vs, nq, in.q, in.sets, tol);
  switch (t->dim) {
    OTG_PICK(2)
    OTG_PICK(3)
    OTG_PICK(4)
    OTG_PICK(5)
    OTG_PICK(6)
  }
#undef OTG_PICK
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

template <int D, bool BOX>
static Result *ray_all_impl(Ctx *ctx, const ViewSet &vs, uint64_t nq, const double *q, const uint32_t *qs, double tol,
                            double inc, double max_dist) {
  Buf<uint32_t> order = query_order<D>(ctx, vs, q, 2 * D, qs, nq);
  const uint32_t *ord = order.p;
  const bool wave = D <= 3 && nq < kRayWaveMaxQueries;  // one wavefront per ray
  Buf<double> dist;
  Result *r = count_fill_csr(
      ctx, nq,
      [&](uint32_t *cnt) {
        if constexpr (D <= 3) {
          if (wave) {
            launch(ctx, "ray_all_count", k_ray_all_wave<D, BOX, false>, dim3((uint32_t)nq), dim3(64), 0, vs.views,
                   vs.qview, q, qs, ord, nq, tol, inc, max_dist, cnt, (const uint64_t *)nullptr, (double *)nullptr,
                   (uint32_t *)nullptr);
            return;
          }
        }
        launch(ctx, "ray_all_count", k_ray_all<D, BOX, false>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views,
               vs.qview, q, qs, ord, nq, tol, inc, max_dist, cnt, (const uint64_t *)nullptr, (double *)nullptr,
               (uint32_t *)nullptr);
      },
      [&](Result *res, uint32_t total) {
        dist = Buf<double>(ctx, std::max<uint32_t>(total, 1));
        if constexpr (D <= 3) {
          if (wave) {
            launch(ctx, "ray_all_fill", k_ray_all_wave<D, BOX, true>, dim3((uint32_t)nq), dim3(64), 0, vs.views,
                   vs.qview, q, qs, ord, nq, tol, inc, max_dist, (uint32_t *)nullptr,
                   (const uint64_t *)res->d_offsets, dist.p, res->d_ids);
            return;
          }
        }
        launch(ctx, "ray_all_fill", k_ray_all<D, BOX, true>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views, vs.qview,
               q, qs, ord, nq, tol, inc, max_dist, (uint32_t *)nullptr, (const uint64_t *)res->d_offsets, dist.p,
               res->d_ids);
      });
  launch(ctx, "ray_all_sort", k_sort_band<false, uint64_t>, dim3(div_up(nq, kQB)), dim3(kQB), 0,
         (const uint64_t *)r->d_offsets, nq, dist.p, r->d_ids);
  return r;
}

Result *query_ray_all(const Tree *t, uint64_t nq, const double *rays, const uint32_t *sets, int mem, double tol,
                      double inc, double max_dist) {
  // the reference tree, as for the first hit (node hit tests decide the walk)
  Ctx *ctx = run_ctx(t);
  DevInput in;
  stage_inputs(ctx, in, rays, sets, nq * 2 * t->dim, nq, mem, (uint32_t)t->sets.size());
  const bool box = t->kind == OTG_KIND_BOX;
  const ViewSet vs = single_view(t);
#define OTG_RAY_ALL(DD)                                                                   \
  case DD:                                                                                \
    return box ? ray_all_impl<DD, true>(ctx, vs, nq, in.q, in.sets, tol, inc, max_dist)   \
               : ray_all_impl<DD, false>(ctx, vs, nq, in.q, in.sets, tol, inc, max_dist);
  switch (t->dim) {
    OTG_RAY_ALL(2)
    OTG_RAY_ALL(3)
    OTG_RAY_ALL(4)
    OTG_RAY_ALL(5)
    OTG_RAY_ALL(6)
  }
#undef OTG_RAY_ALL
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

// Node heap capacity that no ray walk of these trees can outgrow: the node count.
static uint32_t ray_heap_nodes(const ViewSet &vs) {
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(vs.max_nodes, 1), 0x7FFFFFFFull);
}

template <int D, bool BOX>
static Result *ray_impl(Ctx *ctx, const ViewSet &vs, uint64_t nq, const double *q, const uint32_t *qs, double tol,
                        double inc, double max_dist) {
  constexpr int RCAP = 16;
  constexpr uint64_t kScratch = 1ull << 30;  // bytes of global node heaps / candidate lists at a time
  const uint64_t nq1 = std::max<uint64_t>(nq, 1);
  Buf<uint32_t> ids(ctx, nq * RCAP), cnt(ctx, nq1), ovf(ctx, 3), band(ctx, nq1), redo(ctx, nq1);
  Buf<double> Mq(ctx, nq1);
  OTG_HIP_CHECK(hipMemsetAsync(ovf.p, 0, 12, ctx->stream));  // [0] band count, [1] redo count, [2] deep failures
  Buf<uint32_t> order = query_order<D>(ctx, vs, q, 2 * D, qs, nq);
  if (nq < kRayWaveMaxQueries)
    launch(ctx, "ray_first", k_ray_wave<D, BOX, RCAP>, dim3(div_up(nq, kWaveQueries)), dim3(64 * kWaveQueries), 0,
           vs.views, vs.qview, q, qs, (const uint32_t *)order.p, nq, tol, inc, max_dist, ids.p, cnt.p, ovf.p, band.p,
           Mq.p, redo.p, ovf.p + 1);
  else
    launch(ctx, "ray_first", k_ray<D, BOX, RCAP>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views, vs.qview, q, qs,
           (const uint32_t *)order.p, nq, tol, inc, max_dist, ids.p, cnt.p, ovf.p, band.p, Mq.p, redo.p, ovf.p + 1);
  uint32_t hovf[3] = {0, 0, 0};
  OTG_HIP_CHECK(hipMemcpyAsync(hovf, ovf.p, 12, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  const uint32_t hcap = ray_heap_nodes(vs);
  if (hovf[1]) {
    sort_positions(ctx, redo.p, hovf[1], nq);
    Buf<uint32_t> deep_fail(ctx, hovf[1]);
    launch(ctx, "ray_deep", k_ray_deep<D, BOX, RCAP, false>, dim3(div_up(hovf[1], kQB)), dim3(kQB), 0, vs.views,
           vs.qview, q, qs, (const uint32_t *)order.p, (const uint32_t *)redo.p, hovf[1], tol, inc, max_dist, ids.p,
           cnt.p, ovf.p, band.p, Mq.p, ovf.p + 2, deep_fail.p, (uint32_t *)nullptr, (double *)nullptr, 0u);
    OTG_HIP_CHECK(hipMemcpyAsync(hovf, ovf.p, 12, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (hovf[2]) {
      // frontiers beyond the private heap: a node heap of the tree's size in global memory
      const uint32_t nf = hovf[2];
      const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(nf, kScratch / (12ull * hcap)));
      Buf<uint32_t> hn(ctx, chunk * hcap);
      Buf<double> hp(ctx, chunk * hcap);
      Buf<uint32_t> gfail(ctx, chunk + 1);  // [0] count, then the list (a heap of the tree's size never fills)
      OTG_HIP_CHECK(hipMemsetAsync(gfail.p, 0, 4, ctx->stream));
      for (uint64_t c0 = 0; c0 < nf; c0 += chunk) {
        const uint32_t m = (uint32_t)std::min<uint64_t>(chunk, nf - c0);
        launch(ctx, "ray_deep_global", k_ray_deep<D, BOX, RCAP, true>, dim3(div_up(m, kQB)), dim3(kQB), 0, vs.views,
               vs.qview, q, qs, (const uint32_t *)order.p, (const uint32_t *)deep_fail.p + c0, m, tol, inc, max_dist,
               ids.p, cnt.p, ovf.p, band.p, Mq.p, gfail.p, gfail.p + 1, hn.p, hp.p, hcap);
        uint32_t hf = 0;
        OTG_HIP_CHECK(hipMemcpyAsync(&hf, gfail.p, 4, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        if (hf) throw Error{OTG_ERR_HIP, "a ray walk outgrew a node heap of the tree's size"};
      }
      OTG_HIP_CHECK(hipMemcpyAsync(hovf, ovf.p, 12, hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
    }
  }
  const uint32_t nb = hovf[0];
  if (nb == 0) return fixed_to_csr(ctx, ids.p, cnt.p, RCAP, nq);
  // Replay the overflowing rays with candidate lists in global memory: every ray
  // whose list fills is replayed again with one eight times longer (at most the set's
  // entity count, which no ray can exceed), one whose node heap fills with a heap of
  // the tree's node count; lists and heaps are taken a bounded number of rays at a time.
  std::vector<uint32_t> hband(nb);
  OTG_HIP_CHECK(hipMemcpyAsync(hband.data(), band.p, (uint64_t)nb * 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  uint64_t max_n = 1;
  max_n = std::max<uint64_t>(max_n, vs.max_entities);
  struct Done {
    uint32_t q, cnt;
    std::vector<uint32_t> ids;
  };
  std::vector<Done> done;
  done.reserve(nb);
  std::vector<uint32_t> todo = hband;
  uint32_t cap = 256;
  bool global_heap = false;
  Buf<uint32_t> fail(ctx, 1);
  while (!todo.empty()) {
    const uint64_t per_ray = 12ull * cap + (global_heap ? 12ull * hcap : 0ull);
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(todo.size(), kScratch / per_ray));
    std::vector<uint32_t> again_list, again_heap;
    for (uint64_t c0 = 0; c0 < todo.size(); c0 += chunk) {
      const uint32_t m = (uint32_t)std::min<uint64_t>(chunk, todo.size() - c0);
      Buf<uint32_t> lst(ctx, m), bcnt(ctx, m);
      Buf<double> ce(ctx, (uint64_t)m * cap);
      Buf<uint32_t> ci(ctx, (uint64_t)m * cap);
      Buf<uint32_t> hn(ctx, global_heap ? (uint64_t)m * hcap : 0);
      Buf<double> hp(ctx, global_heap ? (uint64_t)m * hcap : 0);
      OTG_HIP_CHECK(hipMemcpyAsync(lst.p, todo.data() + c0, (uint64_t)m * 4, hipMemcpyHostToDevice, ctx->stream));
      OTG_HIP_CHECK(hipMemsetAsync(fail.p, 0, 4, ctx->stream));
      if (global_heap)
        launch(ctx, "ray_replay_global", k_ray_replay<D, BOX, true>, dim3(div_up(m, kQB)), dim3(kQB), 0, vs.views,
               vs.qview, q, qs, (const uint32_t *)lst.p, m, tol, inc, max_dist, cap, ce.p, ci.p, bcnt.p, fail.p, hn.p,
               hp.p, hcap);
      else
        launch(ctx, "ray_replay", k_ray_replay<D, BOX, false>, dim3(div_up(m, kQB)), dim3(kQB), 0, vs.views,
               vs.qview, q, qs, (const uint32_t *)lst.p, m, tol, inc, max_dist, cap, ce.p, ci.p, bcnt.p, fail.p,
               (uint32_t *)nullptr, (double *)nullptr, 0u);
      std::vector<uint32_t> hc(m), hi((uint64_t)m * cap);
      OTG_HIP_CHECK(hipMemcpyAsync(hc.data(), bcnt.p, (uint64_t)m * 4, hipMemcpyDeviceToHost, ctx->stream));
      OTG_HIP_CHECK(hipMemcpyAsync(hi.data(), ci.p, (uint64_t)m * cap * 4, hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
      for (uint32_t i = 0; i < m; ++i) {
        const uint32_t qi = todo[c0 + i];
        if (hc[i] == kReplayListFull) again_list.push_back(qi);
        else if (hc[i] == kReplayHeapFull) again_heap.push_back(qi);
        else done.push_back(Done{qi, hc[i], std::vector<uint32_t>(hi.begin() + (uint64_t)i * cap,
                                                                  hi.begin() + (uint64_t)i * cap + hc[i])});
      }
    }
    if (!again_heap.empty() && global_heap && hcap >= vs.max_nodes)
      throw Error{OTG_ERR_HIP, "a ray walk outgrew a node heap of the tree's size"};
    // heap-full rays first move to the global heap; list-full ones get a longer list
    if (!again_heap.empty()) {
      global_heap = true;
      todo = again_heap;
      todo.insert(todo.end(), again_list.begin(), again_list.end());
      if (!again_list.empty()) cap = (uint32_t)std::min<uint64_t>((uint64_t)cap * 8, max_n);
    } else {
      todo = again_list;
      if (!todo.empty()) {
        if (cap >= max_n) throw Error{OTG_ERR_HIP, "a ray outgrew a candidate list of the set's size"};
        cap = (uint32_t)std::min<uint64_t>((uint64_t)cap * 8, max_n);
      }
    }
  }
  // splice: each replayed ray's count into cnt, then its ids after the CSR is made
  std::vector<uint32_t> dq(done.size()), dc(done.size());
  for (size_t i = 0; i < done.size(); ++i) dq[i] = done[i].q, dc[i] = done[i].cnt;
  const uint32_t nd = (uint32_t)done.size();
  Buf<uint32_t> dqd(ctx, nd), dcd(ctx, nd);
  OTG_HIP_CHECK(hipMemcpyAsync(dqd.p, dq.data(), (uint64_t)nd * 4, hipMemcpyHostToDevice, ctx->stream));
  OTG_HIP_CHECK(hipMemcpyAsync(dcd.p, dc.data(), (uint64_t)nd * 4, hipMemcpyHostToDevice, ctx->stream));
  launch(ctx, "ray_band_counts", k_band_counts, dim3(div_up(nd, 256)), dim3(256), 0, (const uint32_t *)dqd.p, nd,
         (const uint32_t *)dcd.p, cnt.p);
  Result *r = fixed_to_csr(ctx, ids.p, cnt.p, RCAP, nq);
  std::vector<uint64_t> poff(nd + 1, 0);
  for (uint32_t i = 0; i < nd; ++i) poff[i + 1] = poff[i] + done[i].cnt;
  std::vector<uint32_t> packed(std::max<uint64_t>(poff[nd], 1));
  for (uint32_t i = 0; i < nd; ++i) std::copy(done[i].ids.begin(), done[i].ids.end(), packed.begin() + poff[i]);
  Buf<uint64_t> poffd(ctx, nd + 1);
  Buf<uint32_t> packedd(ctx, packed.size());
  OTG_HIP_CHECK(hipMemcpyAsync(poffd.p, poff.data(), (uint64_t)(nd + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
  OTG_HIP_CHECK(hipMemcpyAsync(packedd.p, packed.data(), packed.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  launch(ctx, "ray_replay_splice", k_replay_splice, dim3(div_up(nd, 256)), dim3(256), 0, (const uint32_t *)dqd.p, nd,
         (const uint64_t *)poffd.p, (const uint32_t *)packedd.p, (const uint64_t *)r->d_offsets, r->d_ids);
  ctx->sync();  // the host arrays above must outlive the copies
  return r;
}

static Result *ray_dispatch(Ctx *ctx, int dim, bool box, const ViewSet &vs, uint64_t nq, const double *q,
                            const uint32_t *qs, double tol, double inc, double max_dist) {
#define OTG_RAY(DD)                                                           \
  case DD:                                                                    \
    return box ? ray_impl<DD, true>(ctx, vs, nq, q, qs, tol, inc, max_dist)   \
               : ray_impl<DD, false>(ctx, vs, nq, q, qs, tol, inc, max_dist);
  switch (dim) {
    OTG_RAY(2)
    OTG_RAY(3)
    OTG_RAY(4)
    OTG_RAY(5)
    OTG_RAY(6)
  }
#undef OTG_RAY
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

Result *query_ray(const Tree *t, uint64_t nq, const double *rays, const uint32_t *sets, int mem, double tol, double inc,
                  double max_dist) {
  // Rays always walk the reference tree: each hit whose exit is below the bound
  // assigns M = (1 + inc) exit + tol (ot_query.h:2093-2096), which can raise M, so the
  // result depends on the order hits are met -- the reference tree's order.
  Ctx *ctx = run_ctx(t);
  DevInput in;
  stage_inputs(ctx, in, rays, sets, nq * 2 * t->dim, nq, mem, (uint32_t)t->sets.size());
  ViewSet vs = single_view(t);
  vs.views = reference_device_view(t);  // hit order within a leaf: equal keys by id
  return ray_dispatch(ctx, t->dim, t->kind == OTG_KIND_BOX, vs, nq, in.q, in.sets, tol, inc, max_dist);
}

// One launch over queries on several trees of the same kind and dimension
// (serving).  qview[i] indexes `trees`; qk[i] is query i's k (kNN).  Host inputs.
Result *query_multi(Ctx *ctx, int query, const std::vector<const Tree *> &trees, uint64_t nq, const double *q,
                    const uint32_t *qview, const uint32_t *qk, double p0, double p1, double p2, int mode) {
  if (trees.empty()) throw Error{OTG_ERR_INVALID_ARG, "no trees"};
  const int kind = trees[0]->kind, dim = trees[0]->dim;
  std::vector<TreeView> hv;
  ViewSet vs;
  vs.nviews = (uint32_t)trees.size();
  uint64_t total_n = 0;
  for (const Tree *t : trees) {
    if (t->kind != kind || t->dim != dim) throw Error{OTG_ERR_INVALID_ARG, "trees of one launch differ in kind or dim"};
    const Tree *x = query == 0 ? knn_index(t) : (query == 1 ? range_index(t) : t);
    hv.push_back(query == 2 ? reference_view_of(x) : view_of(x));
    vs.max_nodes = std::max<uint64_t>(vs.max_nodes, x->node_count);
    vs.max_entities = std::max<uint64_t>(vs.max_entities, x->n);
    for (const SetParams &sp : x->sets) vs.max_set = std::max<uint64_t>(vs.max_set, sp.ent_cnt);
    vs.cbits = std::max(vs.cbits, x->layout.cbits);
    vs.sbits = std::max(vs.sbits, x->layout.sbits);
    total_n += x->n;
  }
  uint32_t kmax = 0;
  if (query == 0)
    for (uint64_t i = 0; i < nq; ++i) kmax = std::max(kmax, qk[i]);
  if (kmax > 64) throw Error{OTG_ERR_UNSUPPORTED, "k > 64"};
  for (uint64_t i = 0; i < nq; ++i)
    if (qview[i] >= vs.nviews) throw Error{OTG_ERR_INVALID_ARG, "query tree index"};
  const uint64_t stride = query == 0 ? dim : 2 * dim;
  Buf<TreeView> dv(ctx, hv.size());
  Buf<double> dq(ctx, std::max<uint64_t>(nq * stride, 1));
  Buf<uint32_t> dqv(ctx, std::max<uint64_t>(nq, 1)), dqk(ctx, std::max<uint64_t>(nq, 1));
  OTG_HIP_CHECK(hipMemcpyAsync(dv.p, hv.data(), hv.size() * sizeof(TreeView), hipMemcpyHostToDevice, ctx->stream));
  if (nq) {
    OTG_HIP_CHECK(hipMemcpyAsync(dq.p, q, nq * stride * 8, hipMemcpyHostToDevice, ctx->stream));
    OTG_HIP_CHECK(hipMemcpyAsync(dqv.p, qview, nq * 4, hipMemcpyHostToDevice, ctx->stream));
    if (query == 0) OTG_HIP_CHECK(hipMemcpyAsync(dqk.p, qk, nq * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  vs.views = dv.p;
  vs.qview = dqv.p;
  Result *r;
  if (nq == 0 || total_n == 0 || (query == 0 && kmax == 0)) r = empty_result(ctx, nq);
  else if (query == 0)
    r = knn_dispatch(ctx, dim, kind == OTG_KIND_BOX, vs, nq, dq.p, nullptr, dqk.p, 0, kmax, knn_F0(p0, p1), p1);
  else if (query == 1) r = range_dispatch(ctx, dim, kind == OTG_KIND_BOX, vs, nq, dq.p, nullptr, mode, p0);
  else r = ray_dispatch(ctx, dim, kind == OTG_KIND_BOX, vs, nq, dq.p, nullptr, p0, p1, p2);
  // the staged host arrays must outlive the copies: wait before returning
  ctx->sync();
  return r;
}

// Left set's entities query the right set's tree (SELF: the same set, partners
// with a larger id).  Pairs (left id, right id) packed as left << jbits | right
// and radix-sorted.
template <int D, bool SELF>
static Result *collide_impl(const Tree *lt, uint32_t lset, const Tree *rt, uint32_t rset, double tol) {
  Ctx *ctx = run_ctx(lt);
  const uint64_t ne = lt->sets[lset].ent_cnt;
  const uint64_t nr = rt->sets[rset].ent_cnt;
  TreeView lv = view_of(lt), rv = view_of(rt);
  Result *r = new Result();
  r->ctx = ctx;
  Buf<uint32_t> cnt(ctx, std::max<uint64_t>(ne, 1));
  launch(ctx, "collide_count", k_collide<D, false, SELF>, dim3(div_up(ne, kQB)), dim3(kQB), 0, lv, lset, rv, rset, tol,
         cnt.p, (const uint64_t *)nullptr, (uint64_t *)nullptr);
  Buf<uint32_t> scan(ctx, ne + 1), total(ctx, 1);
  exclusive_scan_u32(ctx, cnt.p, scan.p, ne, total.p);
  uint32_t np = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&np, total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  // the pair list holds 2 ids per pair, indexed with 32-bit positions
  if (np == kScanOverflow || np > 0x7FFFFFFFu) {
    delete r;
    throw Error{OTG_ERR_UNSUPPORTED, "2^31 or more collision pairs"};
  }
  Buf<uint64_t> off(ctx, ne + 1);
  launch(ctx, "offsets_u64", k_offsets_u64, dim3(div_up(ne + 1, 256)), dim3(256), 0, (const uint32_t *)scan.p, ne, np,
         off.p);
  Buf<uint64_t> pairs(ctx, std::max<uint64_t>(np, 1));
  launch(ctx, "collide_fill", k_collide<D, true, SELF>, dim3(div_up(ne, kQB)), dim3(kQB), 0, lv, lset, rv, rset, tol,
         (uint32_t *)nullptr, (const uint64_t *)off.p, pairs.p);
  const uint64_t idmax = std::max(ne, nr);
  const uint32_t jbits = idmax > 1 ? (64 - __builtin_clzll(idmax - 1)) : 1;
  Buf<uint32_t> dummy(ctx, std::max<uint64_t>(np, 1));
  launch(ctx, "pack_pairs", k_pack_pairs, dim3(div_up(np, 256)), dim3(256), 0, pairs.p, (uint64_t)np, jbits);
  radix_sort_pairs(ctx, pairs.p, dummy.p, np, 2 * jbits);
  r->nq = np;
  r->nids = 2ull * np;
  r->d_offsets = (uint64_t *)ctx->alloc((np + 1) * 8);
  r->d_ids = (uint32_t *)ctx->alloc(std::max<uint64_t>(2ull * np, 1) * 4);
  launch(ctx, "split_pairs", k_split_pairs, dim3(div_up(np + 1, 256)), dim3(256), 0, (const uint64_t *)pairs.p,
         (uint64_t)np, jbits, r->d_ids, r->d_offsets);
  return r;
}

template <bool SELF>
static Result *collide_dispatch(const Tree *lt, uint32_t lset, const Tree *rt, uint32_t rset, double tol) {
  switch (lt->dim) {
    case 2: return collide_impl<2, SELF>(lt, lset, rt, rset, tol);
    case 3: return collide_impl<3, SELF>(lt, lset, rt, rset, tol);
    case 4: return collide_impl<4, SELF>(lt, lset, rt, rset, tol);
    case 5: return collide_impl<5, SELF>(lt, lset, rt, rset, tol);
    case 6: return collide_impl<6, SELF>(lt, lset, rt, rset, tol);
  }
  throw Error{OTG_ERR_INVALID_ARG, "dimension"};
}

// ---------------------------------------------------------------------------
// kNN plans: one kNN batch captured as a HIP graph, replayed on new contents of
// the same query buffers
// ---------------------------------------------------------------------------
// Set ids of a replay are checked on the device: an id outside the tree's sets
// reads set 0 instead and raises the plan's error flag (the run then fails).
__global__ void k_sanitize_sets(const uint32_t *__restrict__ in, uint64_t n, uint32_t nsets, uint32_t *__restrict__ out,
                                uint32_t *__restrict__ bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    uint32_t v = in[i];
    if (v >= nsets) {
      v = 0;
      atomicOr(bad, 1u);
    }
    out[i] = v;
  }
}

// CSR offsets with the total read on the device (no host round trip in a graph)
__global__ void k_offsets_dev(const uint32_t *__restrict__ scan, uint64_t n, const uint32_t *__restrict__ total,
                              uint64_t *__restrict__ off) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) off[i] = scan[i];
  if (i == n) off[n] = *total;
}

// The plan's output steps, each a bounded grid that returns at once when its case does
// not hold (flags: [0] band count, [4] a list of another length than k): rows written
// straight into the CSR join the fixed lists, which are then compacted.
__global__ void k_plan_direct_to_fixed(const uint32_t *__restrict__ direct, const uint32_t *__restrict__ cnt,
                                       uint64_t nq, uint32_t k, uint32_t kcap, uint32_t *__restrict__ fixed,
                                       const uint32_t *__restrict__ flags) {
  if (!flags[0] && !flags[4]) return;
  for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < nq * k; x += (uint64_t)gridDim.x * 256) {
    const uint64_t qi = x / k;
    if (cnt[qi] == k) fixed[qi * kcap + (x - qi * k)] = direct[x];
  }
}

__global__ void k_plan_compact(const uint32_t *__restrict__ src, uint32_t stride, const uint64_t *__restrict__ off,
                               uint64_t nq, uint32_t *__restrict__ dst, const uint32_t *__restrict__ flags) {
  if (!flags[4]) return;
  for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < nq * stride; x += (uint64_t)gridDim.x * 256) {
    const uint64_t qi = x / stride;
    const uint32_t t = (uint32_t)(x - qi * stride);
    const uint64_t o = off[qi];
    if (t < off[qi + 1] - o) dst[o + t] = src[x];
  }
}

struct KnnPlan {
  const Tree *t = nullptr;  // the index walked (finer index when the tree has one)
  Ctx *ctx = nullptr;
  uint64_t nq = 0;
  const double *q = nullptr;
  const uint32_t *sets = nullptr;
  uint32_t k = 0, kcap = 0;
  double F0 = 0.0, tol = 0.0;
  Tuning tuning;
  char *mem = nullptr;  // every buffer of the plan, from the context allocator
  // [0] band count, [1] bad set id, [2] id total, [3] persistent tile counter, [4] a list
  // of another length than k (the rows written straight into csr_ids are then moved)
  uint32_t *flags = nullptr;
  uint32_t *h_flags = nullptr;  // pinned copy, written by the graph's last node
  uint32_t *ids_fixed = nullptr, *cnt = nullptr, *band = nullptr, *sets_clean = nullptr, *scan = nullptr;
  uint32_t *csr_ids = nullptr;
  uint64_t *offsets = nullptr;
  double *Fq = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::unique_ptr<Result> band_result;  // the last run's result when a tolerance band outgrew a list
  void (*complete)(KnnPlan &) = nullptr;
  ~KnnPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_flags) (void)hipHostFree(h_flags);
    if (ctx && mem) ctx->free(mem);
  }
};

template <int D, int KCAP, bool BOX>
static void knn_plan_body(KnnPlan &p, Arena &ar) {
  Ctx *ctx = p.ctx;
  const uint64_t nq = p.nq;
  ViewSet vs = single_view(p.t);
  launch(ctx, "fill_u32", k_fill_u32, dim3(1), dim3(64), 0, p.flags, (uint64_t)8, 0u);
  const KnnOut out{p.ids_fixed, p.csr_ids, p.flags + 4};
  const uint32_t *qs = nullptr;
  if (p.sets) {
    launch(ctx, "check_sets", k_sanitize_sets, dim3(std::min<uint64_t>(div_up(nq, 256), 1024)), dim3(256), 0, p.sets,
           nq, (uint32_t)p.t->sets.size(), p.sets_clean, p.flags + 1);
    qs = p.sets_clean;
  }
  Buf<uint32_t> order = query_order<D>(ctx, vs, p.q, D, qs, nq, &ar);
  bool waved = false;
  if constexpr (kKnnWave<D, KCAP, BOX>) {
    if (ctx->tuning.knn_wave && knn_wave_batch<D>(nq)) {
      launch(ctx, "knn", k_knn_wave<D, KCAP>, dim3(div_up(nq, kWaveQueries)), dim3(64 * kWaveQueries), 0, vs.views,
             vs.qview, p.q, qs, (const uint32_t *)nullptr, (const uint32_t *)order.p, nq, p.k, p.F0, p.tol, out, p.cnt,
             p.flags, p.band, p.Fq);
      waved = true;
    }
  }
  if (waved) {
  } else if (ctx->tuning.persistent) {
    const void *kp = (const void *)k_knn_persistent<D, KCAP, BOX>;
    const size_t pad = occupancy_pad(ctx, kp, ctx->tuning.waves_per_simd);
    launch(ctx, "knn", k_knn_persistent<D, KCAP, BOX>, dim3(persistent_grid(ctx, kp, kQB, pad, div_up(nq, kQB))),
           dim3(kQB), pad, vs.views, vs.qview, p.q, qs, (const uint32_t *)nullptr, (const uint32_t *)order.p, nq, p.k,
           p.F0, p.tol, out, p.cnt, p.flags, p.band, p.Fq, p.flags + 3);
  } else {
    const size_t pad = occupancy_pad(ctx, (const void *)k_knn<D, KCAP, BOX>, ctx->tuning.waves_per_simd);
    launch(ctx, "knn", k_knn<D, KCAP, BOX>, dim3(div_up(nq, kQB)), dim3(kQB), pad, vs.views, vs.qview, p.q, qs,
           (const uint32_t *)nullptr, (const uint32_t *)order.p, nq, p.k, p.F0, p.tol, out, p.cnt, p.flags, p.band, p.Fq);
  }
  // every list of k ids went straight into its CSR row (offsets i * k); otherwise the
  // rows join the fixed lists (also for the band pass) and are compacted
  const uint32_t G = (uint32_t)std::min<uint64_t>(div_up(nq * p.k, 256), 16384);
  launch(ctx, "direct_to_fixed", k_plan_direct_to_fixed, dim3(G), dim3(256), 0, (const uint32_t *)p.csr_ids,
         (const uint32_t *)p.cnt, nq, p.k, (uint32_t)KCAP, p.ids_fixed, (const uint32_t *)p.flags);
  exclusive_scan_u32(ctx, p.cnt, p.scan, nq, p.flags + 2, &ar);
  launch(ctx, "offsets_dev", k_offsets_dev, dim3(div_up(nq + 1, 256)), dim3(256), 0, (const uint32_t *)p.scan, nq,
         (const uint32_t *)(p.flags + 2), p.offsets);
  launch(ctx, "compact_fixed", k_plan_compact, dim3((uint32_t)std::min<uint64_t>(div_up(nq * KCAP, 256), 16384)),
         dim3(256), 0, (const uint32_t *)p.ids_fixed, (uint32_t)KCAP, (const uint64_t *)p.offsets, nq, p.csr_ids,
         (const uint32_t *)p.flags);
  OTG_HIP_CHECK(hipMemcpyAsync(p.h_flags, p.flags, 16, hipMemcpyDeviceToHost, ctx->stream));
}

// Queries whose tolerance band outgrew the list: the band pass, outside the graph.
template <int D, int KCAP, bool BOX>
static void knn_plan_complete(KnnPlan &p) {
  Ctx *ctx = p.ctx;
  const uint32_t nb = p.h_flags[0];
  ViewSet vs = single_view(p.t);
  const uint32_t *qs = p.sets ? p.sets_clean : nullptr;
  Buf<uint32_t> bcnt(ctx, nb), boff(ctx, nb + 1), btot(ctx, 1);
  launch(ctx, "knn_band_count", k_knn_band<D, BOX, false>, dim3(div_up(nb, kQB)), dim3(kQB), 0, vs.views, vs.qview,
         p.q, qs, (const uint32_t *)p.band, nb, (const double *)p.Fq, p.tol, bcnt.p, (const uint32_t *)nullptr,
         (double *)nullptr, (uint32_t *)nullptr);
  exclusive_scan_u32(ctx, bcnt.p, boff.p, nb, btot.p);
  uint32_t hb = 0;
  OTG_HIP_CHECK(hipMemcpyAsync(&hb, btot.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  check_id_total(hb);
  OTG_HIP_CHECK(hipMemcpyAsync(boff.p + nb, &hb, 4, hipMemcpyHostToDevice, ctx->stream));
  Buf<double> bd(ctx, std::max<uint32_t>(hb, 1));
  Buf<uint32_t> bi(ctx, std::max<uint32_t>(hb, 1));
  launch(ctx, "knn_band_fill", k_knn_band<D, BOX, true>, dim3(div_up(nb, kQB)), dim3(kQB), 0, vs.views, vs.qview, p.q,
         qs, (const uint32_t *)p.band, nb, (const double *)p.Fq, p.tol, (uint32_t *)nullptr, (const uint32_t *)boff.p,
         bd.p, bi.p);
  launch(ctx, "knn_band_sort", k_sort_band<false, uint32_t>, dim3(div_up(nb, kQB)), dim3(kQB), 0,
         (const uint32_t *)boff.p, (uint64_t)nb, bd.p, bi.p);
  launch(ctx, "knn_band_counts", k_band_counts, dim3(div_up(nb, 256)), dim3(256), 0, (const uint32_t *)p.band, nb,
         (const uint32_t *)bcnt.p, p.cnt);
  std::unique_ptr<Result> r(fixed_to_csr(ctx, p.ids_fixed, p.cnt, KCAP, p.nq));
  launch(ctx, "knn_band_copy", k_band_copy, dim3(div_up(nb, 256)), dim3(256), 0, (const uint32_t *)p.band, nb,
         (const uint32_t *)boff.p, (const uint32_t *)bi.p, (const uint64_t *)r->d_offsets, r->d_ids);
  ctx->sync();
  p.band_result = std::move(r);
}

template <int D, int KCAP, bool BOX>
static void knn_plan_capture(KnnPlan &p) {
  Ctx *ctx = p.ctx;
  const uint64_t nq = p.nq;
  // every buffer up front: the captured work allocates nothing
  const uint64_t ntiles = (nq + 4095) / 4096;
  const size_t bytes = nq * (8ull * KCAP + 112) + ntiles * 2 * 256 * 4 + (8u << 20);
  p.mem = (char *)ctx->alloc(bytes);
  Arena fixed{p.mem, bytes, 0};
  p.flags = (uint32_t *)fixed.take(32);
  p.ids_fixed = (uint32_t *)fixed.take(nq * KCAP * 4);
  p.cnt = (uint32_t *)fixed.take(nq * 4);
  p.band = (uint32_t *)fixed.take(nq * 4);
  p.sets_clean = (uint32_t *)fixed.take(nq * 4);
  p.scan = (uint32_t *)fixed.take(nq * 4 + 4);
  p.csr_ids = (uint32_t *)fixed.take(nq * KCAP * 4);
  p.offsets = (uint64_t *)fixed.take((nq + 1) * 8);
  p.Fq = (double *)fixed.take(nq * 8);
  Arena scratch{fixed.base + fixed.used, fixed.cap - fixed.used, 0};
  OTG_HIP_CHECK(hipHostMalloc(&p.h_flags, 16, hipHostMallocDefault));
  std::memset(p.h_flags, 0, 16);
  // launch-shape queries before the capture (cached)
  if (ctx->tuning.persistent) {
    const void *kp = (const void *)k_knn_persistent<D, KCAP, BOX>;
    (void)persistent_grid(ctx, kp, kQB, occupancy_pad(ctx, kp, ctx->tuning.waves_per_simd), div_up(nq, kQB));
  } else {
    (void)occupancy_pad(ctx, (const void *)k_knn<D, KCAP, BOX>, ctx->tuning.waves_per_simd);
  }
  ctx->sync();
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  OTG_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  try {
    knn_plan_body<D, KCAP, BOX>(p, scratch);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(ctx->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    ctx->profiling = prof;
    throw;
  }
  ctx->profiling = prof;
  OTG_HIP_CHECK(hipStreamEndCapture(ctx->stream, &p.graph));
  OTG_HIP_CHECK(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
  p.complete = &knn_plan_complete<D, KCAP, BOX>;
}

KnnPlan *knn_plan_create(const Tree *tree, uint64_t nq, const double *q, const uint32_t *sets, uint32_t k,
                         double max_within, double tol) {
  if (k == 0 || k > 64) throw Error{OTG_ERR_UNSUPPORTED, "plans take 1 <= k <= 64"};
  if (nq == 0 || nq > 0xFFFFFFFFull / 65) throw Error{OTG_ERR_INVALID_ARG, "plan query count"};
  std::unique_ptr<KnnPlan> p(new KnnPlan());
  p->t = knn_index(tree);
  p->ctx = tree->ctx;
  p->nq = nq;
  p->q = q;
  p->sets = sets;
  p->k = k;
  p->F0 = knn_F0(max_within, tol);
  p->tol = tol;
  p->tuning = p->ctx->tuning;
  if (p->t->n == 0) throw Error{OTG_ERR_INVALID_ARG, "plan on an empty tree"};
  const int dim = p->t->dim;
  const bool box = p->t->kind == OTG_KIND_BOX;
#define OTG_PLAN_K(DD, BB)                                                   \
  if (k <= 8) knn_plan_capture<DD, 9, BB>(*p);                               \
  else if (k <= 16) knn_plan_capture<DD, 17, BB>(*p);                        \
  else if (k <= 32) knn_plan_capture<DD, 33, BB>(*p);                        \
  else knn_plan_capture<DD, 65, BB>(*p);
#define OTG_PLAN(DD)                    \
  case DD:                              \
    if (box) {                          \
      OTG_PLAN_K(DD, true)              \
    } else {                            \
      OTG_PLAN_K(DD, false)             \
    }                                   \
    break;
  switch (dim) {
    OTG_PLAN(2)
    OTG_PLAN(3)
    OTG_PLAN(4)
    OTG_PLAN(5)
    OTG_PLAN(6)
    default:
      throw Error{OTG_ERR_INVALID_ARG, "dimension"};
  }
#undef OTG_PLAN
#undef OTG_PLAN_K
  return p.release();
}

// After the plan's graph ran (alone or inside a larger graph) and the stream drained:
// the checks, and the band pass when a tolerance band outgrew a list.
static void knn_plan_after(KnnPlan *p) {
  if (p->h_flags[1]) throw Error{OTG_ERR_INVALID_ARG, "a query's set id is not a set of the tree"};
  if (p->h_flags[2] == kScanOverflow) throw Error{OTG_ERR_UNSUPPORTED, "a batch result of 2^32 - 1 or more ids"};
  if (p->h_flags[0]) p->complete(*p);
}

void knn_plan_run(KnnPlan *p) {
  Ctx *ctx = p->ctx;
  p->band_result.reset();
  OTG_HIP_CHECK(hipGraphLaunch(p->exec, ctx->stream));
  ctx->sync();
  knn_plan_after(p);
}

// Range plan: one RangeSearch batch on a 2-D / 3-D tree captured as a HIP graph with
// every buffer allocated up front.  The result's id capacity comes from an eager run of
// the batch at creation (ids + ids / 8 + 65,536); deferred runs hold more than
// kInlineCopy ids each and the count walks count a capped query's runs twice, so
// 2 x ids / (kInlineCopy + 1) bounds them.  Counts stay on the device: the capped and
// heavy lists drive bounded grids that stride over them, and a total beyond a capacity
// sets an abort word that every writing kernel checks (the run then reports
// OTG_ERR_CAPACITY and leaves every list empty).  The wavefront count walks keep their
// records for the first keep_q capped queries (the creation run's capped count plus a
// sixteenth and 256), keep_cap records each plus the shared overflow area, as the
// eager path does; capped queries beyond them are filled by the fill walk.
struct RangePlan {
  const Tree *t = nullptr;  // the index walked
  Ctx *ctx = nullptr;
  uint64_t nq = 0;
  const double *q = nullptr;
  const uint32_t *sets = nullptr;
  int mode = 0;
  double tol = 0.0;
  uint64_t cap_ids = 0, cap_copies = 0;
  uint32_t keep_q = 0, keep_cap = 0, keep_x = 0;
  char *mem = nullptr;
  // [0] bad set id, [1] id total, [2] capped queries, [3] heavy queries, [4] runs of the
  // count walks, [5] abort, [6] ids of heavy queries, [7] running run count, [8] queries
  // the fill walk fills, [9] overflow records handed out
  uint32_t *flags = nullptr;
  uint32_t *h_flags = nullptr;
  uint64_t *offsets = nullptr;
  uint32_t *ids = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~RangePlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_flags) (void)hipHostFree(h_flags);
    if (ctx && mem) ctx->free(mem);
  }
};

constexpr uint32_t kPlanListGrid = 16384;  // blocks striding over a device-side query list
constexpr uint32_t kRangePlanFlags = 16;    // device words of a range plan (RangePlan::flags)

// flags[4] = the count walks' run count, then the running count restarts for the fill;
// abort when the ids or the runs outgrow the plan
__global__ void k_range_plan_check(uint32_t *__restrict__ flags, uint64_t cap_ids, uint64_t cap_copies) {
  flags[4] = flags[7];
  flags[7] = 0u;
  flags[5] = ((uint64_t)flags[1] > cap_ids || (uint64_t)flags[4] > cap_copies) ? 1u : 0u;
}

template <int D, bool BOX>
static void range_plan_body(RangePlan &p, Arena &ar) {
  static_assert(D <= 3, "range plans walk 2-D / 3-D trees");
  Ctx *ctx = p.ctx;
  const uint64_t nq = p.nq;
  ViewSet vs = single_view(p.t);
  uint32_t *F = p.flags;
  fill_u32(ctx, F, kRangePlanFlags, 0u);
  const uint32_t *qs = nullptr;
  if (p.sets) {
    Buf<uint32_t> clean(&ar, nq);
    launch(ctx, "check_sets", k_sanitize_sets, dim3(std::min<uint64_t>(div_up(nq, 256), 1024)), dim3(256), 0, p.sets,
           nq, (uint32_t)p.t->sets.size(), clean.p, F + 0);
    qs = clean.p;
  }
  Buf<uint32_t> order = query_order<D>(ctx, vs, p.q, 2 * D, qs, nq, &ar);
  const uint32_t *ord = order.p;
  Buf<uint32_t> cnt(&ar, nq), inl(&ar, nq * kRangeInline), capped(&ar, nq), pos(&ar, nq), clist(&ar, nq);
  Buf<uint32_t> scan(&ar, nq + 1), heavy(&ar, nq), hpos0(&ar, nq), hlist(&ar, nq), hc(&ar, nq), hpos(&ar, nq);
  Buf<RangeCopy> copies(&ar, p.cap_copies);
  Buf<uint64_t> keys(&ar, p.cap_ids);
  Buf<uint32_t> vals(&ar, p.cap_ids);
  Buf<uint32_t> walk(&ar, nq), wpos(&ar, nq), wlist(&ar, nq), keep_ok(&ar, nq), keep_n(&ar, p.keep_q),
      keep_xq(&ar, p.keep_x);
  Buf<RangeRec> keep_recs(&ar, (uint64_t)p.keep_q * p.keep_cap + p.keep_x);
  RangeKeep keep{};
  if (p.keep_q)
    keep = RangeKeep{keep_recs.p, keep_n.p,   keep_ok.p, keep_recs.p + (uint64_t)p.keep_q * p.keep_cap,
                     keep_xq.p,   F + 9,      p.keep_cap, p.keep_x, p.keep_q};
  const uint32_t G = (uint32_t)std::min<uint64_t>(nq, kPlanListGrid);
  // 1. one-lane count walks; the capped ones counted by a wavefront each
  launch(ctx, "range_count", k_range<D, BOX, false>, dim3(div_up(nq, kQB)), dim3(kQB), 0, vs.views, vs.qview, p.q, qs,
         ord, nq, p.mode, p.tol, cnt.p, (const uint64_t *)nullptr, (uint32_t *)nullptr, F + 7, (RangeCopy *)nullptr,
         inl.p, kRangeVisitCap);
  launch(ctx, "range_capped", k_range_capped, dim3(div_up(nq, 256)), dim3(256), 0, ord, nq, (const uint32_t *)cnt.p,
         capped.p);
  exclusive_scan_u32(ctx, capped.p, pos.p, nq, F + 2, &ar);
  launch(ctx, "range_heavy_list", k_range_heavy_list, dim3(div_up(nq, 256)), dim3(256), 0, (const uint32_t *)capped.p,
         (const uint32_t *)pos.p, nq, ord, clist.p);
  launch(ctx, "range_wave_count", k_range_wave<D, BOX, false>, dim3(G), dim3(64), 0, vs.views, vs.qview, p.q, qs,
         (const uint32_t *)clist.p, p.mode, p.tol, cnt.p, (const uint64_t *)nullptr, (uint32_t *)nullptr, F + 7,
         (RangeCopy *)nullptr, (const uint32_t *)(F + 2), (const uint32_t *)nullptr, keep);
  // 2. offsets and the capacity check
  exclusive_scan_u32(ctx, cnt.p, scan.p, nq, F + 1, &ar);
  launch(ctx, "offsets_dev", k_offsets_dev, dim3(div_up(nq + 1, 256)), dim3(256), 0, (const uint32_t *)scan.p, nq,
         (const uint32_t *)(F + 1), p.offsets);
  launch(ctx, "range_plan_check", k_range_plan_check, dim3(1), dim3(1), 0, F, p.cap_ids, p.cap_copies);
  const uint32_t *abort = F + 5;
  // 3. small results from the inline rows; the kept ones from their records; the rest
  // of the heavy ones filled by a wavefront each
  launch(ctx, "range_inline", k_range_inline, dim3(div_up(nq, 256)), dim3(256), 0, ord, nq, (const uint32_t *)cnt.p,
         (const uint32_t *)capped.p, (const uint64_t *)p.offsets, (const uint32_t *)inl.p, p.ids, heavy.p, abort,
         walk.p, (const uint32_t *)(p.keep_q ? keep_ok.p : nullptr));
  exclusive_scan_u32(ctx, heavy.p, hpos0.p, nq, F + 3, &ar);
  launch(ctx, "range_heavy_list", k_range_heavy_list, dim3(div_up(nq, 256)), dim3(256), 0, (const uint32_t *)heavy.p,
         (const uint32_t *)hpos0.p, nq, ord, hlist.p);
  exclusive_scan_u32(ctx, walk.p, wpos.p, nq, F + 8, &ar);
  launch(ctx, "range_heavy_list", k_range_heavy_list, dim3(div_up(nq, 256)), dim3(256), 0, (const uint32_t *)walk.p,
         (const uint32_t *)wpos.p, nq, ord, wlist.p);
  if (p.keep_q)
    launch(ctx, "range_kept", k_range_kept<D, BOX>, dim3(p.keep_q + 256), dim3(256), 0, vs.views, vs.qview, p.q,
           (const uint32_t *)clist.p, p.keep_q, p.mode, p.tol, keep, (const uint64_t *)p.offsets, p.ids,
           (const uint32_t *)(F + 2), abort);
  launch(ctx, "range_fill", k_range_wave<D, BOX, true>, dim3(G), dim3(64), 0, vs.views, vs.qview, p.q, qs,
         (const uint32_t *)wlist.p, p.mode, p.tol, (uint32_t *)nullptr, (const uint64_t *)p.offsets, p.ids, F + 7,
         copies.p, (const uint32_t *)(F + 8), abort, RangeKeep{});
  // 4. long contained runs, one wavefront each
  launch(ctx, "range_copy", k_range_copy, dim3(kPlanListGrid / 4), dim3(256), 0, (const RangeCopy *)copies.p, 0u, p.ids,
         (const uint32_t *)(F + 7), abort);
  // 5. the heavy results sorted by id: one radix sort of (rank in the list, id) keys
  launch(ctx, "range_heavy_counts", k_range_heavy_counts, dim3(div_up(nq, 256)), dim3(256), 0,
         (const uint32_t *)hlist.p, 0u, (const uint64_t *)p.offsets, hc.p, (const uint32_t *)(F + 3));
  exclusive_scan_u32(ctx, hc.p, hpos.p, DevN{nq, F + 3}, F + 6, &ar);
  launch(ctx, "range_sort_keys", k_range_sort_keys, dim3(G), dim3(256), 0, (const uint32_t *)hlist.p, 0u,
         (const uint64_t *)p.offsets, (const uint32_t *)hpos.p, (const uint32_t *)p.ids, keys.p,
         (const uint32_t *)(F + 3), abort, id_bits(vs));
  radix_sort_pairs(ctx, keys.p, vals.p, DevN{p.cap_ids, F + 6},
                   id_bits(vs) + std::max(1u, bit_width32((uint32_t)(nq - 1))), &ar);
  launch(ctx, "range_sort_back", k_range_sort_back, dim3(G), dim3(256), 0, (const uint32_t *)hlist.p, 0u,
         (const uint64_t *)p.offsets, (const uint32_t *)hpos.p, (const uint64_t *)keys.p, p.ids,
         (const uint32_t *)(F + 3), abort, id_bits(vs));
  OTG_HIP_CHECK(hipMemcpyAsync(p.h_flags, F, kRangePlanFlags * 4, hipMemcpyDeviceToHost, ctx->stream));
}

template <int D, bool BOX>
static void range_plan_capture(RangePlan &p) {
  Ctx *ctx = p.ctx;
  const uint64_t nq = p.nq;
  OTG_HIP_CHECK(hipHostMalloc(&p.h_flags, kRangePlanFlags * 4, hipHostMallocDefault));
  std::memset(p.h_flags, 0, kRangePlanFlags * 4);
  const uint64_t tiles = nq / 4096 + 2, id_tiles = p.cap_ids / 4096 + 2;
  size_t bytes = nq * (4 * 18 + 4 * kRangeInline + 8 + 12 + 12) + tiles * 2048 * 2 + p.cap_ids * (4 + 8 + 4 + 12) +
                 id_tiles * 2048 + p.cap_copies * sizeof(RangeCopy) +
                 ((uint64_t)p.keep_q * p.keep_cap + p.keep_x) * sizeof(RangeRec) + (p.keep_q + p.keep_x) * 4ull +
                 (8u << 20);
  for (;; bytes *= 2) {
    p.mem = (char *)ctx->alloc(bytes);
    Arena ar{p.mem, bytes, 0};
    p.flags = (uint32_t *)ar.take(kRangePlanFlags * 4);
    p.offsets = (uint64_t *)ar.take((nq + 1) * 8);
    p.ids = (uint32_t *)ar.take(std::max<uint64_t>(p.cap_ids, 1) * 4);
    ctx->sync();
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    OTG_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    bool short_arena = false;
    try {
      range_plan_body<D, BOX>(p, ar);
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

static void range_plan_after(RangePlan *p) {
  const uint32_t *f = p->h_flags;
  if (f[0]) throw Error{OTG_ERR_INVALID_ARG, "a query's set id is not a set of the tree"};
  if (f[1] == kScanOverflow) throw Error{OTG_ERR_UNSUPPORTED, "a batch result of 2^32 - 1 or more ids"};
  if (f[5]) throw Error{OTG_ERR_CAPACITY, "the batch's results outgrow the plan (make a new plan)"};
}

RangePlan *range_plan_create(const Tree *tree, uint64_t nq, const double *q, const uint32_t *sets, int mode,
                             double tol) {
  if (nq == 0 || nq >= 0x7FFFFFFFull) throw Error{OTG_ERR_INVALID_ARG, "plan query count"};
  if (tree->dim > 3) throw Error{OTG_ERR_UNSUPPORTED, "range plans walk 2-D and 3-D trees"};
  std::unique_ptr<RangePlan> p(new RangePlan());
  p->t = range_index(tree);
  p->ctx = tree->ctx;
  p->nq = nq;
  p->q = q;
  p->sets = sets;
  p->mode = mode;
  p->tol = tol;
  if (p->t->n == 0) throw Error{OTG_ERR_INVALID_ARG, "plan on an empty tree"};
  uint64_t capped = 0;
  {  // the batch's current contents size the plan
    std::unique_ptr<Result> r(query_range(tree, nq, q, sets, OTG_MEM_DEVICE, mode, tol));
    p->cap_ids = r->nids + r->nids / 8 + 65536;
    capped = r->capped;
  }
  if (capped) {  // record slots as the eager path sizes them, for the creation run's capped count and more
    const uint64_t kq = std::min<uint64_t>(nq, capped + capped / 16 + 256);
    const uint32_t want = p->ctx->tuning.range_records;
    if (kq * (want ? want : kKeepMin) <= kKeepBudget) {
      p->keep_q = (uint32_t)kq;
      p->keep_cap = want ? want : (uint32_t)std::min<uint64_t>(kKeepMax, kKeepBudget / kq);
      p->keep_x = (uint32_t)std::min<uint64_t>(kKeepOverflow, 1024ull * p->keep_cap);
    }
  }
  p->cap_copies = 2 * p->cap_ids / (kInlineCopy + 1) + 1024;
  const bool box = p->t->kind == OTG_KIND_BOX;
  if (p->t->dim == 2) {
    if (box) range_plan_capture<2, true>(*p);
    else range_plan_capture<2, false>(*p);
  } else {
    if (box) range_plan_capture<3, true>(*p);
    else range_plan_capture<3, false>(*p);
  }
  return p.release();
}

void range_plan_run(RangePlan *p) {
  Ctx *ctx = p->ctx;
  OTG_HIP_CHECK(hipGraphLaunch(p->exec, ctx->stream));
  ctx->sync();
  range_plan_after(p);
}

void range_plan_result(const RangePlan *p, const uint64_t **off, const uint32_t **ids, uint64_t *nids) {
  *off = p->offsets;
  *ids = p->ids;
  *nids = p->h_flags[5] ? 0 : p->h_flags[1];
}

void range_plan_free(RangePlan *p) { delete p; }
Ctx *range_plan_ctx(RangePlan *p) { return p->ctx; }
uint64_t range_plan_queries(const RangePlan *p) { return p->nq; }

// A build plan and a kNN and/or range plan on its tree as one graph: the build's graph,
// then the query batches' graphs as child graphs that depend on it only.
struct StepPlan {
  BuildPlan *bp = nullptr;
  KnnPlan *kp = nullptr;    // or null
  RangePlan *rp = nullptr;  // or null
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // both batches: the range graph replays on a second stream beside the kNN graph
  hipStream_t side = nullptr;
  hipEvent_t built = nullptr, ranged = nullptr;
  ~StepPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (built) (void)hipEventDestroy(built);
    if (ranged) (void)hipEventDestroy(ranged);
    if (side) (void)hipStreamDestroy(side);
  }
};

StepPlan *step_plan_create(BuildPlan *bp, KnnPlan *kp, RangePlan *rp) {
  const Tree *t = build_plan_tree(bp);
  if (kp && kp->t != knn_index(t)) throw Error{OTG_ERR_INVALID_ARG, "the kNN plan must be made on the build plan's tree"};
  if (rp && rp->t != range_index(t))
    throw Error{OTG_ERR_INVALID_ARG, "the range plan must be made on the build plan's tree"};
  std::unique_ptr<StepPlan> st(new StepPlan());
  st->bp = bp;
  st->kp = kp;
  st->rp = rp;
  OTG_HIP_CHECK(hipGraphCreate(&st->graph, 0));
  // the build first; the kNN and range batches only read the tree and each has its own
  // buffers, so they are two branches that may run side by side
  hipGraphNode_t build = nullptr;
  OTG_HIP_CHECK(hipGraphAddChildGraphNode(&build, st->graph, nullptr, 0, build_plan_graph(bp)));
  for (hipGraph_t g : {kp ? kp->graph : nullptr, rp ? rp->graph : nullptr}) {
    if (!g) continue;
    hipGraphNode_t node = nullptr;
    OTG_HIP_CHECK(hipGraphAddChildGraphNode(&node, st->graph, &build, 1, g));
  }
  OTG_HIP_CHECK(hipGraphInstantiate(&st->exec, st->graph, nullptr, nullptr, 0));
  if (kp && rp) {
    OTG_HIP_CHECK(hipStreamCreateWithFlags(&st->side, hipStreamNonBlocking));
    OTG_HIP_CHECK(hipEventCreateWithFlags(&st->built, hipEventDisableTiming));
    OTG_HIP_CHECK(hipEventCreateWithFlags(&st->ranged, hipEventDisableTiming));
  }
  return st.release();
}

void step_plan_run(StepPlan *st) {
  Ctx *ctx = build_plan_tree(st->bp)->ctx;
  if (st->kp) st->kp->band_result.reset();
  if (st->side) {
    // build, then kNN on the context's stream and range on the side stream at once
    OTG_HIP_CHECK(hipGraphLaunch(build_plan_exec(st->bp), ctx->stream));
    OTG_HIP_CHECK(hipEventRecord(st->built, ctx->stream));
    OTG_HIP_CHECK(hipStreamWaitEvent(st->side, st->built, 0));
    OTG_HIP_CHECK(hipGraphLaunch(st->rp->exec, st->side));
    OTG_HIP_CHECK(hipGraphLaunch(st->kp->exec, ctx->stream));
    OTG_HIP_CHECK(hipEventRecord(st->ranged, st->side));
    OTG_HIP_CHECK(hipStreamWaitEvent(ctx->stream, st->ranged, 0));
  } else {
    OTG_HIP_CHECK(hipGraphLaunch(st->exec, ctx->stream));
  }
  ctx->sync();
  build_plan_finish(st->bp);
  if (st->kp) knn_plan_after(st->kp);
  if (st->rp) range_plan_after(st->rp);
}

void step_plan_free(StepPlan *st) { delete st; }

Ctx *step_plan_ctx(StepPlan *st) { return build_plan_tree(st->bp)->ctx; }

void knn_plan_result(const KnnPlan *p, const uint64_t **off, const uint32_t **ids, uint64_t *nids) {
  if (p->band_result) {
    *off = p->band_result->d_offsets;
    *ids = p->band_result->d_ids;
    *nids = p->band_result->nids;
  } else {
    *off = p->offsets;
    *ids = p->csr_ids;
    *nids = p->h_flags[2];
  }
}

void knn_plan_free(KnnPlan *p) { delete p; }

Ctx *knn_plan_ctx(KnnPlan *p) { return p->ctx; }

uint64_t knn_plan_queries(const KnnPlan *p) { return p->nq; }

Result *query_collide(const Tree *t, uint32_t set, double tol) {
  if (t->kind != OTG_KIND_BOX) throw Error{OTG_ERR_UNSUPPORTED, "collision detection of point trees"};
  if (set >= t->sets.size()) throw Error{OTG_ERR_INVALID_ARG, "set index"};
  return collide_dispatch<true>(t, set, t, set, tol);
}

Result *query_collide_trees(const Tree *lt, uint32_t lset, const Tree *rt, uint32_t rset, double tol) {
  if (lt->kind != OTG_KIND_BOX || rt->kind != OTG_KIND_BOX)
    throw Error{OTG_ERR_UNSUPPORTED, "collision detection of point trees"};
  if (lt->dim != rt->dim) throw Error{OTG_ERR_INVALID_ARG, "trees of different dimension"};
  if (lset >= lt->sets.size() || rset >= rt->sets.size()) throw Error{OTG_ERR_INVALID_ARG, "set index"};
  if (lt->ctx != rt->ctx && lt->ctx->device != rt->ctx->device)
    throw Error{OTG_ERR_INVALID_ARG, "trees on different devices"};
  return collide_dispatch<false>(lt, lset, rt, rset, tol);
}

}  // namespace otg
