// otg_tune.hip -- the autotuner of the batched build + query path.
//
// It measures one workload (a forest of point or box sets and a batch of kNN and/or
// range queries on it) under candidate settings and keeps the fastest:
//   leaf capacity   {8, 16, 32, 64}   the build's maxElementNumInNode
//                                     (ot_static_linear_core.h:523-529);
//   index capacity  {leaf, 16, 32, 64} leaf capacity of the finer kNN / range index;
//   waves per SIMD  {registers, 2, 1}  occupancy cap of the kNN walks;
//   query order     {Morton, input}    walk order of large batches;
//   kNN launch      {one tile per block, persistent tile pulling}.
// Each later stage keeps the earlier stages' choices.  The pick is cached per
// workload shape (kind, dimension, sets, entities, queries, k) for the process, and
// left in the context's settings, where every later build and query on it uses it.
// Results never depend on the settings: kNN and range ids are functions of the
// entities' distances and boxes only.
#include <cfloat>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "otg_internal.hpp"

namespace otg {
namespace {

using ShapeKey = std::tuple<int, int, uint32_t, uint64_t, uint64_t, uint32_t, uint64_t>;
std::mutex g_mu;
std::map<ShapeKey, otg_tuning> g_cache;

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace otg

extern "C" {

OTG_API otg_status otg_ctx_set_tuning(otg_ctx *ctx, const otg_tuning *t) {
  if (!ctx || !t) return OTG_ERR_INVALID_ARG;
  if (t->waves_per_simd > 8 || t->order_queries > 1 || t->persistent > 1) return OTG_ERR_INVALID_ARG;
  ctx->tuning.index_capacity = t->index_capacity;
  ctx->tuning.waves_per_simd = t->waves_per_simd;
  ctx->tuning.order_queries = t->order_queries;
  ctx->tuning.persistent = t->persistent;
  ctx->tuning.range_records = t->range_records;
  return OTG_OK;
}

OTG_API otg_status otg_ctx_get_tuning(const otg_ctx *ctx, otg_tuning *t) {
  if (!ctx || !t) return OTG_ERR_INVALID_ARG;
  t->leaf_capacity = 0;
  t->index_capacity = ctx->tuning.index_capacity;
  t->waves_per_simd = ctx->tuning.waves_per_simd;
  t->order_queries = ctx->tuning.order_queries;
  t->persistent = ctx->tuning.persistent;
  t->range_records = ctx->tuning.range_records;
  t->step_ms = 0.0;
  return OTG_OK;
}

OTG_API otg_status otg_autotune(otg_ctx *ctx, int kind, int dim, uint32_t nsets, const uint64_t *set_offsets,
                                const double *geom, int mem, const double *boxes_or_null,
                                const uint32_t *max_depths_or_null, uint64_t nq_knn, const double *knn_points,
                                const uint32_t *knn_sets, uint32_t k, uint64_t nq_range, const double *ranges,
                                const uint32_t *range_sets, int qmem, otg_tuning *out) {
  using namespace otg;
  if (!ctx || !out || !set_offsets || nsets == 0) return OTG_ERR_INVALID_ARG;
  if (kind != OTG_KIND_POINT && kind != OTG_KIND_BOX) return OTG_ERR_INVALID_ARG;
  if ((nq_knn && (!knn_points || k == 0)) || (nq_range && !ranges)) return OTG_ERR_INVALID_ARG;
  if (mem != OTG_MEM_HOST && mem != OTG_MEM_DEVICE) return OTG_ERR_INVALID_ARG;
  if (qmem != OTG_MEM_HOST && qmem != OTG_MEM_DEVICE) return OTG_ERR_INVALID_ARG;
  const ShapeKey key{kind, dim, nsets, set_offsets[nsets], nq_knn, k, nq_range};
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      *out = it->second;
      return otg_ctx_set_tuning(ctx, out);
    }
  }
  const Tuning saved = ctx->tuning;
  try {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    // one step: build all sets with capacity `cap`, then the query batches
    auto step = [&](uint32_t cap, bool with_build, Tree *reuse, std::unique_ptr<otg_tree> &keep) -> double {
      std::vector<uint32_t> caps(nsets, cap);
      const double t0 = now_ms();
      if (with_build) {
        keep.reset(new otg_tree());
        keep->ctx = ctx;
        keep->kind = kind;
        keep->dim = dim;
        build_forest(keep.get(), geom, mem, set_offsets, nsets, boxes_or_null, max_depths_or_null, caps.data());
        reuse = keep.get();
      }
      if (nq_knn) {
        std::unique_ptr<Result> r(query_knn(reuse, nq_knn, knn_points, knn_sets, qmem, k, DBL_MAX, kBaseTol));
      }
      if (nq_range) {
        std::unique_ptr<Result> r(query_range(reuse, nq_range, ranges, range_sets, qmem, OTG_RANGE_INSIDE, kBaseTol));
      }
      ctx->sync();
      return now_ms() - t0;
    };
    auto best_of = [&](auto &&run) {
      run();  // warm: allocator cache, code objects
      return std::min(run(), run());
    };
    ctx->tuning = Tuning{};
    std::unique_ptr<otg_tree> tree;
    // 1. leaf capacity (build + queries)
    uint32_t best_cap = 16;
    double best = 1e300;
    for (uint32_t cap : {8u, 16u, 32u, 64u}) {
      const double ms = best_of([&] { return step(cap, true, nullptr, tree); });
      if (ms < best) best = ms, best_cap = cap;
    }
    // 2. finer index capacity (point trees; build + queries)
    Tuning pick{};
    if (kind == OTG_KIND_POINT)
      for (uint32_t icap : {16u, 32u, 64u}) {
        if (icap == best_cap) continue;
        ctx->tuning = pick;
        ctx->tuning.index_capacity = icap;
        const double ms = best_of([&] { return step(best_cap, true, nullptr, tree); });
        if (ms < best) best = ms, pick.index_capacity = icap;
      }
    // 3. occupancy and order (queries only, on the tree of the settings so far)
    ctx->tuning = pick;
    step(best_cap, true, nullptr, tree);
    double qbest = 1e300;
    Tuning qpick = pick;
    for (uint32_t waves : {0u, 2u, 1u})
      for (uint32_t order : {1u, 0u})
        for (uint32_t persist : {0u, 1u}) {
          ctx->tuning = pick;
          ctx->tuning.waves_per_simd = waves;
          ctx->tuning.order_queries = order;
          ctx->tuning.persistent = persist;
          std::unique_ptr<otg_tree> none;
          const double ms = best_of([&] { return step(best_cap, false, tree.get(), none); });
          if (ms < qbest) qbest = ms, qpick = ctx->tuning;
        }
    const double qdefault = [&] {
      ctx->tuning = pick;
      std::unique_ptr<otg_tree> none;
      return best_of([&] { return step(best_cap, false, tree.get(), none); });
    }();
    tree.reset();
    otg_tuning res;
    res.leaf_capacity = best_cap;
    res.index_capacity = qpick.index_capacity;
    res.waves_per_simd = qpick.waves_per_simd;
    res.order_queries = qpick.order_queries;
    res.persistent = qpick.persistent;
    res.range_records = qpick.range_records;
    res.step_ms = best - qdefault + qbest;
    ctx->tuning = qpick;
    {
      std::lock_guard<std::mutex> g(g_mu);
      g_cache[key] = res;
    }
    *out = res;
    ctx->last_error.clear();
    return OTG_OK;
  } catch (const Error &e) {
    ctx->tuning = saved;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->last_error = e.msg;
    return e.status;
  } catch (const std::bad_alloc &) {
    ctx->tuning = saved;
    ctx->last_error = "host allocation failed";
    return OTG_ERR_OOM;
  } catch (...) {
    ctx->tuning = saved;
    ctx->last_error = "unknown error";
    return OTG_ERR_HIP;
  }
}

}  // extern "C"
