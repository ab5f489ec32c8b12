// otg_internal.hpp -- shared host/device declarations of libotg.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/otg/otg.h"

namespace otg {

constexpr int kMaxDim = 6;      // dimensions instantiated on the device
constexpr int kMaxDepth = 19;   // configuration.h:78 MAX_ALLOWED_DEPTH_ID
constexpr double kBaseTol = 2.220446049250313080847e-16 * 100.0;  // general.h:57

// Per-set parameters of a forest (a plain tree is a forest of one set).
struct SetParams {
  uint64_t ent_off;    // first entity (storage position == input position)
  uint64_t ent_cnt;
  uint64_t node_off;   // first node in the forest node table
  uint64_t node_cnt;
  uint32_t max_depth;  // M
  uint32_t cap;        // max elements per node
  uint32_t root_splits;  // filled by the build
  uint32_t min_lcp;      // 1 if a boundary inside the set has level 0, else 255
  double space_min[kMaxDim];
  double space_max[kMaxDim];
  double size[kMaxDim];
  double raster[kMaxDim];
  double deraster[kMaxDim];
  double depth_size[(kMaxDepth + 2) * kMaxDim];  // m_nodeSize per depth
};

// Key layout of the forest sort: set << (cbits + dbits) | code << dbits | depth
struct KeyLayout {
  uint32_t dbits;  // depth bits (boxes) or 0 (points)
  uint32_t cbits;  // D * max(M)
  uint32_t sbits;  // bit width of (nsets - 1)
  uint32_t total;  // sbits + cbits + dbits
};

// Node record used by every traversal.
struct NodeInfo {
  uint32_t ebegin;  // own entity segment (storage positions)
  uint32_t ecount;
  uint32_t cbegin;  // children (global node ids)
  uint32_t ccount;
};

class Ctx;

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// Settings the autotuner chooses (otg_autotune); defaults are the untuned path.
struct Tuning {
  uint32_t index_capacity = 0;  // leaf capacity of the finer kNN / range index (0: the tree's)
  uint32_t waves_per_simd = 0;  // occupancy cap of the kNN walks (0: as registers allow)
  uint32_t order_queries = 1;   // Morton-order batches of >= 16,384 queries before walking
  uint32_t persistent = 0;      // kNN walks as a persistent launch pulling 64-query tiles
  uint32_t range_records = 0;   // kept records per capped RangeSearch query (0: from the batch)
  uint32_t knn_wave = 1;        // 6-D point kNN with k <= 63: one wavefront per query
};

class Ctx {
 public:
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  bool profiling = false;
  Tuning tuning;
  struct KTime {
    double ms = 0;
    uint64_t launches = 0;
  };
  std::map<std::string, KTime> ktimes;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending_events;
  std::vector<hipEvent_t> event_pool;

  // caching allocator: size class -> free blocks
  std::multimap<size_t, void *> free_blocks;
  std::map<void *, size_t> live_blocks;
  uint64_t bytes_in_use = 0, bytes_cached = 0;
  std::mutex mu;

  void *alloc(size_t bytes);
  void free(void *p);
  void release_cache();

  // Error word kernels set when a bounded device-wide wait runs out, and its pinned
  // host copy.  sync() waits for the stream and raises such an error.
  uint32_t *d_err = nullptr;
  uint32_t *h_err = nullptr;
  void sync();

  hipEvent_t get_event();
  void begin_kernel(const char *name, hipEvent_t *start);
  void end_kernel(const char *name, hipEvent_t start);
  void collect_times();  // synchronises the stream
};

// Error carrying a status code across the host orchestration (never crosses the ABI).
struct Error {
  otg_status status;
  std::string msg;
};

// Bump allocator over a fixed block of device memory the caller owns (the
// allocation-free entry point's scratch array).  Buffers taken from it are
// released all at once by the caller, never one by one.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  bool exhausted = false;  // a take did not fit
  void *take(size_t bytes) {
    const size_t b = (bytes + 255) & ~size_t(255);
    if (b > cap - used) {
      exhausted = true;
      throw Error{OTG_ERR_UNSUPPORTED, "the batch needs more scratch memory than the limit"};
    }
    void *p = base + used;
    used += b;
    return p;
  }
};

// Scoped device buffer from the context allocator, or (owning nothing) from an arena.
template <typename T>
struct Buf {
  Ctx *ctx = nullptr;
  T *p = nullptr;
  size_t n = 0;
  Buf() = default;
  Buf(Ctx *c, size_t count) : ctx(c), n(count) { p = count ? (T *)c->alloc(count * sizeof(T)) : nullptr; }
  Buf(Arena *a, size_t count) : ctx(nullptr), n(count) { p = count ? (T *)a->take(count * sizeof(T)) : nullptr; }
  // from the arena when there is one, else from the context
  Buf(Ctx *c, Arena *a, size_t count) : Buf(a ? Buf(a, count) : Buf(c, count)) {}
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  Buf(Buf &&o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  Buf &operator=(Buf &&o) noexcept {
    reset();
    ctx = o.ctx; p = o.p; n = o.n;
    o.p = nullptr; o.n = 0;
    return *this;
  }
  ~Buf() { reset(); }
  void reset() {
    if (p && ctx) ctx->free(p);
    p = nullptr;
    n = 0;
  }
  T *release() { T *q = p; p = nullptr; n = 0; return q; }
};

#define OTG_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t e__ = (expr);                                                             \
    if (e__ != hipSuccess)                                                               \
      throw ::otg::Error{e__ == hipErrorOutOfMemory ? OTG_ERR_OOM : OTG_ERR_HIP,         \
                         std::string(#expr) + ": " + hipGetErrorString(e__)};            \
  } while (0)

// Launch with optional event timing.
template <typename Kernel, typename... Args>
inline void launch(Ctx *ctx, const char *name, Kernel k, dim3 grid, dim3 block, size_t shmem, Args... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipEvent_t ev0 = nullptr;
  if (ctx->profiling) ctx->begin_kernel(name, &ev0);
  hipLaunchKernelGGL(k, grid, block, shmem, ctx->stream, args...);
  OTG_HIP_CHECK(hipGetLastError());
  if (ctx->profiling) ctx->end_kernel(name, ev0);
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// A count the kernels read on the device.  A build captured into a graph knows its
// sizes only there: buffers and grids are sized by `cap`, and the kernels take
// min(*p, cap) elements (a count beyond cap is flagged by the caller).  p null: cap
// is the count.
struct DevN {
  uint64_t cap = 0;
  const uint32_t *p = nullptr;
  __device__ __forceinline__ uint64_t get() const {
    if (!p) return cap;
    const uint64_t v = *p;
    return v < cap ? v : cap;
  }
};

// ---- primitives (otg_primitives.hip) ----
// Exclusive prefix sum of n uint32 values into out (may alias in), accumulated in 64 bits.
// *d_total (device, if not null) receives the total, or kScanOverflow when it does not
// fit 32 bits (out is then not meaningful).
constexpr uint32_t kScanOverflow = 0xFFFFFFFFu;
void exclusive_scan_u32(Ctx *ctx, const uint32_t *in, uint32_t *out, DevN n, uint32_t *d_total, Arena *arena);
inline void exclusive_scan_u32(Ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *d_total,
                               Arena *arena = nullptr) {
  exclusive_scan_u32(ctx, in, out, DevN{n, nullptr}, d_total, arena);
}
// a[0..n) = v by a kernel (no memset node when captured into a graph)
void fill_u32(Ctx *ctx, uint32_t *a, uint64_t n, uint32_t v);
// Stable LSD radix sort of (key, val) pairs on bits [0, bits).  Buffers are swapped internally; the
// result is in keys/vals on return.
void radix_sort_pairs(Ctx *ctx, uint64_t *keys, uint32_t *vals, DevN n, uint32_t bits, Arena *arena);
// Stable LSD radix sort inside each set: the pairs of set s sit in [d_off[s], d_off[s+1])
// (device offsets, nsets + 1 of them) and are sorted there on bits [0, bits) -- the bits
// above (the set id of forest keys) are already in order.  The pairs start in (ka, va);
// (kb, vb) is the other buffer of the ping-pong, so the result ends in (ka, va) after an
// even number of passes ((bits + 7) / 8) and in (kb, vb) after an odd one.
// radix_sets_fit says whether the per-set digit tables are small enough and the sets
// few enough against the tile count for the per-set sort to pay.
bool radix_sets_fit(uint64_t n, uint32_t nsets, uint32_t bits);
// Packed mode (ibits > 0, bits + ibits <= 64): ka holds words (key bits << ibits | index
// of the point inside its set), va is not read, and the last pass writes the keys
// (set << set_shift | key bits; set_shift 64: none) and the global indices.
void radix_sort_sets(Ctx *ctx, uint64_t *ka, uint32_t *va, uint64_t *kb, uint32_t *vb, uint64_t n,
                     const uint64_t *d_off, uint32_t nsets, uint32_t bits, Arena *arena, uint32_t ibits = 0,
                     uint32_t set_shift = 64);
inline void radix_sort_pairs(Ctx *ctx, uint64_t *keys, uint32_t *vals, uint64_t n, uint32_t bits,
                             Arena *arena = nullptr) {
  radix_sort_pairs(ctx, keys, vals, DevN{n, nullptr}, bits, arena);
}
// Boundary-level scans: for v[0..n) (uint8 levels, 0 = hard boundary) computes
// left[i] = set of running minima of v[i], v[i-1], ... (inclusive, stops at level 0) and
// right[i] = same going right, as bit masks.  need (null: every position): only the
// positions x with need[x + 1] (left) or need[x - 1] (right) set are written.
void level_scans(Ctx *ctx, const uint8_t *v, uint64_t n, uint32_t *left_mask, uint32_t *right_mask,
                 const uint8_t *need = nullptr, Arena *arena = nullptr);

// Device-side view of one tree, passed to traversal kernels (an array of them
// lets one launch serve queries on many trees).
struct TreeView {
  const NodeInfo *info;
  const uint32_t *rend;
  const double *nmin;
  const double *nsize;
  const double *geom;  // SoA, sorted order
  uint64_t n;
  const uint32_t *ids;
  const SetParams *sets;
};

// ---- tree ----
struct Tree {
  Ctx *ctx = nullptr;
  int kind = 0;
  int dim = 0;
  uint64_t n = 0;
  uint64_t node_count = 0;
  uint64_t seg_count = 0;  // leaf segments of the last plain node-table build (sizes a build plan)
  std::vector<SetParams> sets;
  KeyLayout layout{};
  SetParams *d_sets = nullptr;
  // entity data in storage (sorted) order
  uint32_t *d_ids = nullptr;      // local entity ids
  mutable uint64_t *d_keys = nullptr;  // sorted keys (a shared sort's reference tree: made on first use)
  double *d_geom = nullptr;       // SoA sorted geometry: comp*n + j; boxes: 2*dim comps (min..., max...)
  // nodes
  NodeInfo *d_info = nullptr;
  uint32_t *d_rend = nullptr;     // subtree entity range end
  uint8_t *d_depth = nullptr;
  uint64_t *d_nkey = nullptr;     // node Morton key
  double *d_nmin = nullptr;       // node_count x dim
  double *d_nsize = nullptr;      // node_count x dim
  uint64_t device_bytes = 0;
  // Finer tree over the same points, used only by kNN when the reference tree has
  // leaves over capacity at max depth (kNN ids depend on distances only).
  Tree *knn_tree = nullptr;
  TreeView *d_view = nullptr;  // device copy of this tree's TreeView
  // A finer index sorted together with its reference tree shares that tree's entity
  // storage (d_ids, d_geom), which the reference tree owns.
  bool shares_entities = false;
  // Arrays in a build plan's arena (the plan frees them, not the tree).
  bool plan_arrays = false;
  // The reference tree of a shared sort keeps equal location keys in the finer
  // index's order instead of by id.  Ray-first (whose bound depends on the order
  // hits are met), the export and the archive use the reference's order (equal keys
  // by id), made on first use and kept: ids, geometry and a device view of them.
  bool fine_order = false;
  mutable std::mutex ref_mu;
  // makes d_keys from the finer index's keys when first needed (export, reference order)
  void (*lazy_keys)(const Tree *) = nullptr;
  mutable uint32_t *d_ids_ref = nullptr;
  mutable double *d_geom_ref = nullptr;
  mutable TreeView *d_view_ref = nullptr;

  ~Tree();
};

// Host TreeView of a built tree, and its device copy in t->d_view (and the finer index's).
TreeView view_of(const Tree *t);
void publish_views(Tree *t);
// The tree's view with its entity storage in the reference's order (equal keys by
// id): the tree's own view unless fine_order, else made on first use.
TreeView reference_view_of(const Tree *t);
const TreeView *reference_device_view(const Tree *t);

// Create<SeqExec> of one tree (kind and dim set): the node numbering and entity order of
// the reference's unsorted build (Build<false>, ot_static_linear_core.h:417-519).
void build_seq(Tree *t, const double *geom, int mem, uint64_t n, const double *box, uint32_t depth, uint32_t cap);
// Loads a StaticLinearOrthoTreeCore archive (as otg_tree_serialize writes it) into t
// (kind and dim set): geom = the n entities in id order (host or device per mem), or
// null for an archive that carries them after the core (OrthoTreeManaged).
void load_static_archive(Tree *t, const uint8_t *buf, uint64_t size, bool large, uint64_t n, const double *geom,
                         int mem);
void build_forest(Tree *t, const double *geom, int mem, const uint64_t *offsets, uint32_t nsets,
                  const double *boxes, const uint32_t *depths, const uint32_t *caps);

// Build plan (point forests): one build captured as a HIP graph with every buffer
// and capacity fixed when it is made (from a plain build of the contents then), and
// replayed on new contents of the same device geometry.  run() synchronises and
// brings the host side of the plan's tree up to date; finish() does that part alone
// after the plan's graph ran inside a larger graph.
struct BuildPlan;
BuildPlan *build_plan_create(Ctx *ctx, int dim, const double *geom, const uint64_t *offsets, uint32_t nsets,
                             const double *boxes, const uint32_t *depths, const uint32_t *caps);
void build_plan_run(BuildPlan *p);
void build_plan_finish(BuildPlan *p);
const Tree *build_plan_tree(const BuildPlan *p);
hipGraph_t build_plan_graph(const BuildPlan *p);
// The tree's sorted keys, made now if they are made on first use (Tree::lazy_keys).
void ensure_keys(const Tree *t);
hipGraphExec_t build_plan_exec(const BuildPlan *p);
void build_plan_free(BuildPlan *p);
// A build plan followed by a kNN plan on its tree, replayed as one graph; run()
// synchronises, then brings the tree's host side up to date and completes the kNN
// batch (band pass) as each plan's own run would.
struct KnnPlan;
struct RangePlan;
struct StepPlan;
StepPlan *step_plan_create(BuildPlan *bp, KnnPlan *kp, RangePlan *rp);
void step_plan_run(StepPlan *st);
void step_plan_free(StepPlan *st);
Ctx *step_plan_ctx(StepPlan *st);

struct Result {
  Ctx *ctx = nullptr;
  uint64_t nq = 0;
  uint64_t nids = 0;
  uint64_t *d_offsets = nullptr;
  uint32_t *d_ids = nullptr;
  uint64_t capped = 0;  // RangeSearch: queries the wavefront walks counted (sizes range plans)
  // host-resident results (served requests): used when host_only is set
  bool host_only = false;
  std::vector<uint64_t> h_offsets;
  std::vector<uint32_t> h_ids;
  ~Result();
};

// Queries run on the calling thread's context when one is set (serving workers).
Ctx *run_ctx(const Tree *t);
void set_thread_run_ctx(Ctx *c);

Result *query_knn(const Tree *t, uint64_t nq, const double *pts, const uint32_t *sets, int mem, uint32_t k,
                  double max_within, double tol,
                  uint64_t *footprint = nullptr);
// Caller-owned CSR of the allocation-free kNN path: offsets[nq + 1] and up to
// `capacity` ids, device memory; nids is set to the result's id count (also when the
// capacity is short, so the caller learns the size it needs).
struct CsrOut {
  uint64_t *offsets = nullptr;
  uint32_t *ids = nullptr;
  uint64_t capacity = 0;
  uint64_t nids = 0;
};
// kNN of device-resident queries with every temporary taken from `arena`: no device
// allocation, queries walked in input order; synchronises before returning.
void query_knn_noalloc(const Tree *t, uint64_t nq, const double *pts, const uint32_t *sets, uint32_t k,
                       double max_within, double tol, Arena &arena, CsrOut &out);
// kNN plan: one batch of device-resident queries captured once as a HIP graph (set
// check, query ordering, walk, CSR) and replayed on new contents of the same
// buffers.  run() synchronises; a tolerance band longer than a list is completed by
// the band pass outside the graph.  The result stays valid until the next run.
struct KnnPlan;
KnnPlan *knn_plan_create(const Tree *t, uint64_t nq, const double *q, const uint32_t *sets, uint32_t k,
                         double max_within, double tol);
void knn_plan_run(KnnPlan *p);
void knn_plan_result(const KnnPlan *p, const uint64_t **offsets, const uint32_t **ids, uint64_t *nids);
void knn_plan_free(KnnPlan *p);
Ctx *knn_plan_ctx(KnnPlan *p);
uint64_t knn_plan_queries(const KnnPlan *p);
Result *query_range(const Tree *t, uint64_t nq, const double *ranges, const uint32_t *sets, int mem, int mode,
                    double tol);
// Range plan (2-D / 3-D trees): one RangeSearch batch of device-resident boxes captured
// as a HIP graph, buffers sized by the batch's result at creation; replays on new
// contents of the same buffers; a result that outgrows the plan fails the run with
// OTG_ERR_CAPACITY.  The result stays valid until the next run.
RangePlan *range_plan_create(const Tree *t, uint64_t nq, const double *q, const uint32_t *sets, int mode, double tol);
void range_plan_run(RangePlan *p);
void range_plan_result(const RangePlan *p, const uint64_t **offsets, const uint32_t **ids, uint64_t *nids);
void range_plan_free(RangePlan *p);
Ctx *range_plan_ctx(RangePlan *p);
uint64_t range_plan_queries(const RangePlan *p);
Result *query_ray(const Tree *t, uint64_t nq, const double *rays, const uint32_t *sets, int mem, double tol,
                  double inc, double max_dist);
// One launch over queries on several trees of one kind and dimension; host inputs.
// query: 0 kNN (qk per-query k, p0 max_within, p1 tol), 1 range (mode, p0 tol),
// 2 ray first (p0 tol, p1 increment, p2 max distance).
Result *query_multi(Ctx *ctx, int query, const std::vector<const Tree *> &trees, uint64_t nq, const double *q,
                    const uint32_t *qview, const uint32_t *qk, double p0, double p1, double p2, int mode);
Result *query_collide(const Tree *t, uint32_t set, double tol);
Result *query_collide_trees(const Tree *lt, uint32_t lset, const Tree *rt, uint32_t rset, double tol);
Result *query_pick(const Tree *t, uint64_t nq, const double *picks, const uint32_t *sets, int mem, double tol);
Result *query_planes(const Tree *t, uint64_t nq, uint32_t nplanes, const double *planes, const uint32_t *sets, int mem,
                     int mode, double tol);
Result *query_ray_all(const Tree *t, uint64_t nq, const double *rays, const uint32_t *sets, int mem, double tol,
                      double inc, double max_dist);

// Looks up the allocation-free entry point's scratch array on `device` (current),
// once per device; a failure is kept and reported by that entry point.
void noalloc_init_device(int device);

// host-side mirrors of reference helpers
uint32_t max_theoretical_depth(int dim);
uint32_t estimate_max_depth(uint64_t n, uint64_t cap, int dim);

}  // namespace otg

// Opaque ABI handles are the internal objects (no extra members).
struct otg_ctx : public otg::Ctx {};
struct otg_tree : public otg::Tree {};
struct otg_result : public otg::Result {};
