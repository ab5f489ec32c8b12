// otg_abi.hip -- the extern "C" boundary of libotg: contexts, trees, results.
// No C++ exception crosses this boundary; every failure becomes a status code.
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <memory>

#include "otg_internal.hpp"

namespace otg {

// ---- caching allocator ----------------------------------------------------
static size_t round_size(size_t b) {
  if (b <= 256) return 256;
  // round to 1/8 of the next power of two to bound waste
  size_t p = 1;
  while (p < b) p <<= 1;
  size_t step = p >> 3;
  return (b + step - 1) / step * step;
}

void *Ctx::alloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  size_t sz = round_size(bytes);
  std::lock_guard<std::mutex> g(mu);
  auto it = free_blocks.lower_bound(sz);
  if (it != free_blocks.end() && it->first <= sz + sz / 4) {
    void *p = it->second;
    size_t have = it->first;
    free_blocks.erase(it);
    live_blocks[p] = have;
    bytes_cached -= have;
    bytes_in_use += have;
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // release the cache and retry once
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
    bytes_cached = 0;
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error{OTG_ERR_OOM, "hipMalloc of " + std::to_string(sz) + " bytes failed"};
    }
  }
  live_blocks[p] = sz;
  bytes_in_use += sz;
  return p;
}

void Ctx::free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu);
  auto it = live_blocks.find(p);
  if (it == live_blocks.end()) return;
  size_t sz = it->second;
  live_blocks.erase(it);
  bytes_in_use -= sz;
  // blocks are reused by later work on the same stream (stream order keeps it safe)
  free_blocks.emplace(sz, p);
  bytes_cached += sz;
}

void Ctx::sync() {
  if (d_err) OTG_HIP_CHECK(hipMemcpyAsync(h_err, d_err, 4, hipMemcpyDeviceToHost, stream));
  OTG_HIP_CHECK(hipStreamSynchronize(stream));
  if (h_err && *h_err) {
    *h_err = 0;
    OTG_HIP_CHECK(hipMemsetAsync(d_err, 0, 4, stream));
    OTG_HIP_CHECK(hipStreamSynchronize(stream));
    throw Error{OTG_ERR_HIP, "a device-wide wait ran past its bound (results discarded)"};
  }
}

void Ctx::release_cache() {
  std::lock_guard<std::mutex> g(mu);
  (void)hipStreamSynchronize(stream);
  for (auto &kv : free_blocks) (void)hipFree(kv.second);
  free_blocks.clear();
  bytes_cached = 0;
}

hipEvent_t Ctx::get_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  OTG_HIP_CHECK(hipEventCreate(&e));
  return e;
}

void Ctx::begin_kernel(const char *name, hipEvent_t *start) {
  *start = get_event();
  OTG_HIP_CHECK(hipEventRecord(*start, stream));
}

void Ctx::end_kernel(const char *name, hipEvent_t start) {
  hipEvent_t stop = get_event();
  OTG_HIP_CHECK(hipEventRecord(stop, stream));
  pending_events.push_back({name, {start, stop}});
}

void Ctx::collect_times() {
  if (pending_events.empty()) return;
  OTG_HIP_CHECK(hipStreamSynchronize(stream));
  for (auto &pe : pending_events) {
    float ms = 0.f;
    OTG_HIP_CHECK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
    KTime &kt = ktimes[pe.first];
    kt.ms += ms;
    kt.launches += 1;
    event_pool.push_back(pe.second.first);
    event_pool.push_back(pe.second.second);
  }
  pending_events.clear();
}

Tree::~Tree() {
  delete knn_tree;
  if (!ctx) return;
  ctx->free(d_view);
  ctx->free(d_ids_ref);
  ctx->free(d_geom_ref);
  ctx->free(d_view_ref);
  if (plan_arrays) return;
  ctx->free(d_sets);
  if (!shares_entities) {
    ctx->free(d_ids);
    ctx->free(d_geom);
  }
  ctx->free(d_keys);
  ctx->free(d_info);
  ctx->free(d_rend);
  ctx->free(d_depth);
  ctx->free(d_nkey);
  ctx->free(d_nmin);
  ctx->free(d_nsize);
}

// ---- static-core archive ----------------------------------------------------
// Byte sink of the reference's BinaryOutputArchive (serialization/binary_archive.h:
// 85-175): arithmetic values in their own width, little-endian; counts as
// serialized_size_t (uint32, or uint64 with ORTHOTREE__LARGE_DATASET).  With a
// NULL destination it only counts.
static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the archive is written in host byte order");

struct ArchiveOut {
  uint8_t *dst;
  bool large;
  uint64_t n = 0;
  template <typename T>
  void put(T v) {
    if (dst) std::memcpy(dst + n, &v, sizeof(T));
    n += sizeof(T);
  }
  void count(uint64_t c) {  // make_size_tag(size) of a std::vector / std::array / variant index
    if (large) {
      put<uint64_t>(c);
    } else {
      if (c > 0xFFFFFFFFull) throw Error{OTG_ERR_UNSUPPORTED, "archive count exceeds serialized_size_t (uint32)"};
      put<uint32_t>((uint32_t)c);
    }
  }
  void vec(const double *v, int D) {  // std::array<double, D> (stl/array.h)
    count((uint64_t)D);
    for (int d = 0; d < D; ++d) put<double>(v[d]);
  }
};

struct ArchiveTable {
  int D;
  bool box;
  uint32_t max_depth, cap;
  uint64_t ne, nn;
  const double *space_min, *space_size, *depth_size;  // depth_size: (max_depth + 2) x D
  const uint8_t *depth;
  const uint32_t *ebegin, *ecount, *cbegin, *ccount, *ids;
  const double *nmin, *nsize;
};

// GetMaxPossibleNodeCount + IsFit (ot_static_linear_core.h:330-344, :367-372):
// 0 = NodeStorage256, 1 = NodeStorage65536, 2 = NodeStorageGeneral.
static int node_storage_index(uint64_t n, uint32_t M, int D) {
  auto full_tree = [D](uint64_t depth) -> uint64_t {  // (2^(D*depth) - 1) / (2^D - 1)
    if ((uint64_t)D * depth >= 64) return ~0ull;
    return ((1ull << (D * depth)) - 1) / ((1ull << D) - 1);
  };
  const uint64_t by_depth = full_tree(M);
  const uint64_t bw = n > 1 ? 64 - (uint64_t)__builtin_clzll(n - 1) : 0;
  const uint64_t full = std::min<uint64_t>(M, bw / (uint64_t)D);
  const uint64_t by_entities = full_tree(full) + ((uint64_t)M - full) * n;
  const uint64_t max_nodes = std::min(by_entities, by_depth);
  if (max_nodes < 0xFFull && n < 0xFFull) return 0;
  if (max_nodes < 0xFFFFull && n < 0xFFFFull) return 1;
  return 2;
}

template <typename Index, typename Seg>
static void write_node_storage(ArchiveOut &ar, const ArchiveTable &t) {
  // nodeChildSegmentBegins starts as {1}; a node with children extends it to
  // id + 2 entries and adds its child count to the last (CreateNode, :383-412)
  std::vector<uint64_t> begins{1};
  for (uint64_t i = 0; i < t.nn; ++i) {
    if (!t.ccount[i]) continue;
    if (begins.back() != t.cbegin[i]) throw Error{OTG_ERR_HIP, "node table children are not in BFS order"};
    begins.resize(i + 2, begins.back());
    begins.back() += t.ccount[i];
  }
  ar.count(begins.size());
  for (uint64_t b : begins) ar.put<Index>((Index)b);
  ar.count(t.nn);
  for (uint64_t i = 0; i < t.nn; ++i) {  // Segment: begin, length (serialization.h:84-89)
    ar.put<Seg>((Seg)t.ebegin[i]);
    ar.put<Seg>((Seg)t.ecount[i]);
  }
}

// serialize(StaticLinearOrthoTreeCore) serialization.h:373-386 and of the base
// OrthoTreeCoreBase :359-371 with the members of ot_base.h:76-89 and
// ot_static_linear_core.h:85-89.  Point trees store MinPoint geometry (no tree
// box, m_nodeSize per depth); loose box trees store MBRs (the nominal tree box,
// no m_nodeSize).  An empty tree keeps the default variant (NodeStorage256,
// empty vectors): Create returns before choosing one (:538-539).
static void write_static_archive(ArchiveOut &ar, const ArchiveTable &t) {
  const int D = t.D;
  ar.put<uint32_t>(0);  // StaticLinearOrthoTreeCore::SERIALIZED_VERSION_ID
  ar.put<uint32_t>(0);  // OrthoTreeCoreBase::SERIALIZED_VERSION_ID
  ar.put<uint64_t>(t.cap);        // m_maxElementNum (std::size_t)
  ar.put<uint32_t>(t.max_depth);  // m_maxDepthID (depth_t)
  if (t.box) {                    // m_nominalTreeBox {minPoint, size}
    ar.vec(t.space_min, D);
    ar.vec(t.space_size, D);
  } else {  // m_nodeSize: max_depth + 2 entries (InitBase, ot_base.h:119-130)
    ar.count(t.max_depth + 2);
    for (uint32_t k = 0; k < t.max_depth + 2; ++k) ar.vec(t.depth_size + (size_t)k * D, D);
  }
  const int kind = t.ne ? node_storage_index(t.ne, t.max_depth, D) : 0;
  ar.count((uint64_t)kind);  // std::variant index (stl/variant.h)
  if (t.ne == 0) {
    ar.count(0);
    ar.count(0);
  } else if (kind == 0) {
    write_node_storage<uint8_t, uint8_t>(ar, t);
  } else if (kind == 1) {
    write_node_storage<uint16_t, uint16_t>(ar, t);
  } else if (ar.large) {
    write_node_storage<uint32_t, uint64_t>(ar, t);
  } else {
    write_node_storage<uint32_t, uint32_t>(ar, t);
  }
  ar.count(t.nn);  // m_nodeDepthIDs
  for (uint64_t i = 0; i < t.nn; ++i) ar.put<uint8_t>(t.depth[i]);
  ar.count(t.ne);  // m_entityStorage
  for (uint64_t i = 0; i < t.ne; ++i) ar.put<uint32_t>(t.ids[i]);
  ar.count(t.nn);  // m_nodeGeometry
  for (uint64_t i = 0; i < t.nn; ++i) {
    ar.vec(t.nmin + i * D, D);
    if (t.box) ar.vec(t.nsize + i * D, D);
  }
}

}  // namespace otg

using otg::Ctx;
using otg::Error;


namespace {

template <typename F>
otg_status guarded(Ctx *ctx, F &&f) {
  try {
    f();
    if (ctx) ctx->last_error.clear();
    return OTG_OK;
  } catch (const Error &e) {
    if (ctx) ctx->last_error = e.msg;
    return e.status;
  } catch (const std::bad_alloc &) {
    if (ctx) ctx->last_error = "host allocation failed";
    return OTG_ERR_OOM;
  } catch (...) {
    if (ctx) ctx->last_error = "unknown error";
    return OTG_ERR_HIP;
  }
}

}  // namespace

extern "C" {

OTG_API const char *otg_status_string(otg_status s) {
  switch (s) {
    case OTG_OK: return "ok";
    case OTG_ERR_INVALID_ARG: return "invalid argument";
    case OTG_ERR_HIP: return "HIP runtime error";
    case OTG_ERR_OOM: return "out of device memory";
    case OTG_ERR_UNSUPPORTED: return "unsupported input";
    case OTG_ERR_OUT_OF_SPACE: return "point outside the given tree box";
    case OTG_ERR_KEY_OVERFLOW: return "key bits exceed 64";
    case OTG_ERR_NO_DEVICE: return "no HIP device";
    case OTG_ERR_CAPACITY: return "output buffer too small";
  }
  return "unknown status";
}

OTG_API otg_status otg_ctx_create(int device, otg_ctx **out) {
  if (!out) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return OTG_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) return OTG_ERR_INVALID_ARG;
  otg_ctx *c = new (std::nothrow) otg_ctx();
  if (!c) return OTG_ERR_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return OTG_ERR_HIP;
  }
  if (hipMalloc(&c->d_err, 256) != hipSuccess || hipMemset(c->d_err, 0, 256) != hipSuccess ||
      hipHostMalloc(&c->h_err, 4, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    if (c->d_err) (void)hipFree(c->d_err);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return OTG_ERR_HIP;
  }
  *c->h_err = 0;
  // the allocation-free entry point's scratch address on this device (once per device)
  otg::noalloc_init_device(device);
  *out = c;
  return OTG_OK;
}

OTG_API void otg_ctx_destroy(otg_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &pe : ctx->pending_events) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto &kv : ctx->free_blocks) (void)hipFree(kv.second);
  for (auto &kv : ctx->live_blocks) (void)hipFree(kv.first);
  if (ctx->d_err) (void)hipFree(ctx->d_err);
  if (ctx->h_err) (void)hipHostFree(ctx->h_err);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

OTG_API otg_status otg_ctx_synchronize(otg_ctx *ctx) {
  if (!ctx) return OTG_ERR_INVALID_ARG;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    OTG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

OTG_API const char *otg_ctx_last_error(const otg_ctx *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

OTG_API otg_status otg_ctx_mem_stats(const otg_ctx *ctx, uint64_t *in_use, uint64_t *cached) {
  if (!ctx) return OTG_ERR_INVALID_ARG;
  if (in_use) *in_use = ctx->bytes_in_use;
  if (cached) *cached = ctx->bytes_cached;
  return OTG_OK;
}

OTG_API otg_status otg_ctx_release_cache(otg_ctx *ctx) {
  if (!ctx) return OTG_ERR_INVALID_ARG;
  return guarded(ctx, [&] { ctx->release_cache(); });
}

OTG_API otg_status otg_ctx_set_profiling(otg_ctx *ctx, int enable) {
  if (!ctx) return OTG_ERR_INVALID_ARG;
  ctx->profiling = enable != 0;
  return OTG_OK;
}

OTG_API otg_status otg_ctx_kernel_times(otg_ctx *ctx, uint32_t cap, char *names, double *total_ms, uint64_t *launches,
                                        uint32_t *count) {
  if (!ctx) return OTG_ERR_INVALID_ARG;
  return guarded(ctx, [&] {
    ctx->collect_times();
    uint32_t i = 0;
    for (auto &kv : ctx->ktimes) {
      if (i < cap) {
        if (names) {
          std::snprintf(names + (size_t)i * 64, 64, "%s", kv.first.c_str());
        }
        if (total_ms) total_ms[i] = kv.second.ms;
        if (launches) launches[i] = kv.second.launches;
      }
      ++i;
    }
    if (count) *count = i;
  });
}

OTG_API otg_status otg_ctx_reset_kernel_times(otg_ctx *ctx) {
  if (!ctx) return OTG_ERR_INVALID_ARG;
  return guarded(ctx, [&] {
    ctx->collect_times();
    ctx->ktimes.clear();
  });
}

OTG_API otg_status otg_forest_build(otg_ctx *ctx, int kind, int dim, uint32_t nsets, const uint64_t *set_offsets,
                                    const double *geom, int mem, const double *boxes_or_null,
                                    const uint32_t *max_depths_or_null, const uint32_t *max_elements, otg_tree **out) {
  if (!ctx || !out || !set_offsets || !max_elements || nsets == 0) return OTG_ERR_INVALID_ARG;
  if (kind != OTG_KIND_POINT && kind != OTG_KIND_BOX) return OTG_ERR_INVALID_ARG;
  if (dim < 2 || dim > otg::kMaxDim) return OTG_ERR_INVALID_ARG;
  if (mem != OTG_MEM_HOST && mem != OTG_MEM_DEVICE) return OTG_ERR_INVALID_ARG;
  for (uint32_t s = 0; s < nsets; ++s)
    if (set_offsets[s + 1] < set_offsets[s]) return OTG_ERR_INVALID_ARG;
  if (set_offsets[nsets] && !geom) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  otg_tree *t = new (std::nothrow) otg_tree();
  if (!t) return OTG_ERR_OOM;
  t->ctx = ctx;
  t->kind = kind;
  t->dim = dim;
  otg_status st = guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    otg::build_forest(t, geom, mem, set_offsets, nsets, boxes_or_null, max_depths_or_null, max_elements);
  });
  if (st != OTG_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    delete t;
    return st;
  }
  *out = t;
  return OTG_OK;
}

OTG_API otg_status otg_tree_build_seq(otg_ctx *ctx, int kind, int dim, uint64_t n, const double *geom, int mem,
                                      const double *box_or_null, uint32_t max_depth, uint32_t max_elements,
                                      otg_tree **out) {
  if (!ctx || !out || max_elements == 0) return OTG_ERR_INVALID_ARG;
  if (kind != OTG_KIND_POINT && kind != OTG_KIND_BOX) return OTG_ERR_INVALID_ARG;
  if (dim < 2 || dim > otg::kMaxDim) return OTG_ERR_INVALID_ARG;
  if (mem != OTG_MEM_HOST && mem != OTG_MEM_DEVICE) return OTG_ERR_INVALID_ARG;
  if (n && !geom) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  otg_tree *t = new (std::nothrow) otg_tree();
  if (!t) return OTG_ERR_OOM;
  t->ctx = ctx;
  t->kind = kind;
  t->dim = dim;
  otg_status st = guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    otg::build_seq(t, geom, mem, n, box_or_null, max_depth, max_elements);
  });
  if (st != OTG_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    delete t;
    return st;
  }
  *out = t;
  return OTG_OK;
}

OTG_API otg_status otg_tree_deserialize(otg_ctx *ctx, int kind, int dim, const uint8_t *buf, uint64_t size,
                                        uint32_t flags, uint64_t n, const double *geom, int mem, otg_tree **out) {
  if (!ctx || !out || !buf || (flags & ~OTG_ARCHIVE_LARGE_DATASET)) return OTG_ERR_INVALID_ARG;
  if (kind != OTG_KIND_POINT && kind != OTG_KIND_BOX) return OTG_ERR_INVALID_ARG;
  if (dim < 2 || dim > otg::kMaxDim) return OTG_ERR_INVALID_ARG;
  if (geom && mem != OTG_MEM_HOST && mem != OTG_MEM_DEVICE) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  otg_tree *t = new (std::nothrow) otg_tree();
  if (!t) return OTG_ERR_OOM;
  t->ctx = ctx;
  t->kind = kind;
  t->dim = dim;
  otg_status st = guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    otg::load_static_archive(t, buf, size, (flags & OTG_ARCHIVE_LARGE_DATASET) != 0, n, geom, mem);
  });
  if (st != OTG_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    delete t;
    return st;
  }
  *out = t;
  return OTG_OK;
}

OTG_API otg_status otg_tree_build(otg_ctx *ctx, int kind, int dim, uint64_t n, const double *geom, int mem,
                                  const double *box_or_null, uint32_t max_depth, uint32_t max_elements, otg_tree **out) {
  uint64_t offs[2] = {0, n};
  return otg_forest_build(ctx, kind, dim, 1, offs, geom, mem, box_or_null, max_depth ? &max_depth : nullptr,
                          &max_elements, out);
}

OTG_API void otg_tree_free(otg_tree *t) {
  if (!t) return;
  if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
  delete t;
}

OTG_API otg_status otg_tree_get_info(const otg_tree *t, otg_tree_info *out) {
  if (!t || !out) return OTG_ERR_INVALID_ARG;
  out->kind = t->kind;
  out->dim = t->dim;
  out->nsets = (uint32_t)t->sets.size();
  out->reserved = 0;
  out->entity_count = t->n;
  out->node_count = t->node_count;
  out->device_bytes = t->device_bytes;
  return OTG_OK;
}

OTG_API otg_status otg_tree_get_set_info(const otg_tree *t, uint32_t set, otg_set_info *out) {
  if (!t || !out || set >= t->sets.size()) return OTG_ERR_INVALID_ARG;
  const otg::SetParams &sp = t->sets[set];
  out->entity_offset = sp.ent_off;
  out->entity_count = sp.ent_cnt;
  out->node_offset = sp.node_off;
  out->node_count = sp.node_cnt;
  out->max_depth = sp.max_depth;
  out->max_elements = sp.cap;
  for (int d = 0; d < 16; ++d) {
    out->space_min[d] = d < t->dim ? sp.space_min[d] : 0.0;
    out->space_max[d] = d < t->dim ? sp.space_max[d] : 0.0;
  }
  return OTG_OK;
}

OTG_API otg_status otg_tree_export(const otg_tree *t, uint32_t set, uint8_t *node_depth, uint32_t *entity_begin,
                                   uint32_t *entity_count, uint32_t *child_begin, uint32_t *child_count,
                                   uint64_t *node_key, double *node_min, double *node_size, uint32_t *entity_ids,
                                   uint64_t *sorted_keys) {
  if (!t || set >= t->sets.size()) return OTG_ERR_INVALID_ARG;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    const otg::SetParams &sp = t->sets[set];
    const uint64_t nn = sp.node_cnt, no = sp.node_off, ne = sp.ent_cnt, eo = sp.ent_off;
    const int D = t->dim;
    hipStream_t st = ctx->stream;
    std::vector<otg::NodeInfo> info(nn);
    if (nn) OTG_HIP_CHECK(hipMemcpyAsync(info.data(), t->d_info + no, nn * sizeof(otg::NodeInfo), hipMemcpyDeviceToHost, st));
    if (node_depth && nn) OTG_HIP_CHECK(hipMemcpyAsync(node_depth, t->d_depth + no, nn, hipMemcpyDeviceToHost, st));
    if (node_key && nn) OTG_HIP_CHECK(hipMemcpyAsync(node_key, t->d_nkey + no, nn * 8, hipMemcpyDeviceToHost, st));
    if (node_min && nn) OTG_HIP_CHECK(hipMemcpyAsync(node_min, t->d_nmin + no * D, nn * D * 8, hipMemcpyDeviceToHost, st));
    if (node_size && nn) OTG_HIP_CHECK(hipMemcpyAsync(node_size, t->d_nsize + no * D, nn * D * 8, hipMemcpyDeviceToHost, st));
    if (entity_ids && ne) {
      const uint32_t *ids = otg::reference_view_of(t).ids;
      OTG_HIP_CHECK(hipMemcpyAsync(entity_ids, ids + eo, ne * 4, hipMemcpyDeviceToHost, st));
    }
    std::vector<uint64_t> keys;
    if (sorted_keys && ne) otg::ensure_keys(t);
    if (sorted_keys && ne && !t->d_keys)
      throw Error{OTG_ERR_UNSUPPORTED, "a point tree loaded from an archive has no location keys (the archive holds none)"};
    if (sorted_keys && ne) {
      keys.resize(ne);
      OTG_HIP_CHECK(hipMemcpyAsync(keys.data(), t->d_keys + eo, ne * 8, hipMemcpyDeviceToHost, st));
    }
    OTG_HIP_CHECK(hipStreamSynchronize(st));
    for (uint64_t i = 0; i < nn; ++i) {
      if (entity_begin) entity_begin[i] = (uint32_t)(info[i].ebegin - eo);
      if (entity_count) entity_count[i] = info[i].ecount;
      if (child_begin) child_begin[i] = info[i].ccount ? (uint32_t)(info[i].cbegin - no) : 0u;
      if (child_count) child_count[i] = info[i].ccount;
    }
    if (sorted_keys) {
      const otg::KeyLayout &L = t->layout;
      const uint64_t mask = (L.cbits + L.dbits >= 64) ? ~0ull : ((1ull << (L.cbits + L.dbits)) - 1);
      for (uint64_t i = 0; i < ne; ++i) sorted_keys[i] = keys[i] & mask;
    }
  });
}

OTG_API otg_status otg_tree_serialize(const otg_tree *t, uint32_t set, uint32_t flags, uint8_t *buf,
                                      uint64_t buf_size, uint64_t *size) {
  if (!t || !size || set >= t->sets.size() || (flags & ~OTG_ARCHIVE_LARGE_DATASET)) return OTG_ERR_INVALID_ARG;
  const otg::SetParams &sp = t->sets[set];
  const uint64_t nn = sp.node_cnt, ne = sp.ent_cnt;
  std::vector<uint8_t> depth(nn);
  std::vector<uint32_t> eb(nn), ec(nn), cb(nn), cc(nn), ids(ne);
  std::vector<double> nmin(nn * t->dim), nsize(nn * t->dim);
  otg_status st = otg_tree_export(t, set, depth.data(), eb.data(), ec.data(), cb.data(), cc.data(), nullptr,
                                  nmin.data(), nsize.data(), ids.data(), nullptr);
  if (st != OTG_OK) return st;
  return guarded(t->ctx, [&] {
    otg::ArchiveTable tab{t->dim, t->kind == OTG_KIND_BOX, sp.max_depth, sp.cap, ne, nn, sp.space_min, sp.size,
                          sp.depth_size, depth.data(), eb.data(), ec.data(), cb.data(), cc.data(), ids.data(),
                          nmin.data(), nsize.data()};
    const bool large = (flags & OTG_ARCHIVE_LARGE_DATASET) != 0;
    otg::ArchiveOut probe{nullptr, large};
    otg::write_static_archive(probe, tab);
    *size = probe.n;
    if (!buf) return;
    if (buf_size < probe.n) throw Error{OTG_ERR_INVALID_ARG, "archive buffer is smaller than the archive"};
    otg::ArchiveOut out{buf, large};
    otg::write_static_archive(out, tab);
  });
}

OTG_API otg_status otg_knn(const otg_tree *t, uint64_t nq, const double *points, const uint32_t *set_ids, int mem,
                           uint32_t k, double max_within, double tolerance, otg_result **out) {
  if (!t || !out || (nq && !points)) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(otg::query_knn(t, nq, points, set_ids, mem, k, max_within, tolerance));
  });
}

OTG_API otg_status otg_knn_plan_create(const otg_tree *t, uint64_t nq, const double *points, const uint32_t *set_ids,
                                       uint32_t k, double max_within, double tolerance, otg_knn_plan **out) {
  if (!t || !out || !points || nq == 0) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = reinterpret_cast<otg_knn_plan *>(otg::knn_plan_create(t, nq, points, set_ids, k, max_within, tolerance));
  });
}

OTG_API otg_status otg_knn_plan_run(otg_knn_plan *plan) {
  if (!plan) return OTG_ERR_INVALID_ARG;
  auto *p = reinterpret_cast<otg::KnnPlan *>(plan);
  Ctx *ctx = otg::knn_plan_ctx(p);
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    otg::knn_plan_run(p);
  });
}

OTG_API otg_status otg_build_plan_create(otg_ctx *ctx, int dim, uint32_t nsets, const uint64_t *set_offsets,
                                         const double *points, const double *boxes, const uint32_t *max_depths_or_null,
                                         const uint32_t *max_elements, otg_build_plan **out) {
  if (!ctx || !out || !set_offsets || !max_elements || !points || !boxes || nsets == 0) return OTG_ERR_INVALID_ARG;
  if (dim < 2 || dim > otg::kMaxDim) return OTG_ERR_INVALID_ARG;
  for (uint32_t s = 0; s < nsets; ++s)
    if (set_offsets[s + 1] < set_offsets[s]) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = reinterpret_cast<otg_build_plan *>(
        otg::build_plan_create(ctx, dim, points, set_offsets, nsets, boxes, max_depths_or_null, max_elements));
  });
}

OTG_API otg_status otg_build_plan_run(otg_build_plan *plan) {
  if (!plan) return OTG_ERR_INVALID_ARG;
  auto *p = reinterpret_cast<otg::BuildPlan *>(plan);
  Ctx *ctx = otg::build_plan_tree(p)->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    otg::build_plan_run(p);
  });
}

OTG_API otg_status otg_build_plan_tree(const otg_build_plan *plan, const otg_tree **tree) {
  if (!plan || !tree) return OTG_ERR_INVALID_ARG;
  *tree = static_cast<const otg_tree *>(otg::build_plan_tree(reinterpret_cast<const otg::BuildPlan *>(plan)));
  return OTG_OK;
}

OTG_API void otg_build_plan_free(otg_build_plan *plan) {
  if (!plan) return;
  auto *p = reinterpret_cast<otg::BuildPlan *>(plan);
  Ctx *ctx = otg::build_plan_tree(p)->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  otg::build_plan_free(p);
}

OTG_API otg_status otg_step_plan_create(otg_build_plan *build, otg_knn_plan *knn_or_null,
                                        otg_range_plan *range_or_null, otg_step_plan **out) {
  if (!build || !out) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  auto *bp = reinterpret_cast<otg::BuildPlan *>(build);
  Ctx *ctx = otg::build_plan_tree(bp)->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = reinterpret_cast<otg_step_plan *>(otg::step_plan_create(bp, reinterpret_cast<otg::KnnPlan *>(knn_or_null),
                                                                   reinterpret_cast<otg::RangePlan *>(range_or_null)));
  });
}

OTG_API otg_status otg_range_plan_create(const otg_tree *t, uint64_t nq, const double *ranges, const uint32_t *set_ids,
                                         int mode, double tolerance, otg_range_plan **out) {
  if (!t || !out || !ranges || nq == 0) return OTG_ERR_INVALID_ARG;
  if (mode != OTG_RANGE_INSIDE && mode != OTG_RANGE_OVERLAP) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = reinterpret_cast<otg_range_plan *>(otg::range_plan_create(t, nq, ranges, set_ids, mode, tolerance));
  });
}

OTG_API otg_status otg_range_plan_run(otg_range_plan *plan) {
  if (!plan) return OTG_ERR_INVALID_ARG;
  auto *p = reinterpret_cast<otg::RangePlan *>(plan);
  Ctx *ctx = otg::range_plan_ctx(p);
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    otg::range_plan_run(p);
  });
}

OTG_API otg_status otg_range_plan_result(const otg_range_plan *plan, const uint64_t **offsets, const uint32_t **ids,
                                         uint64_t *nids) {
  if (!plan || !offsets || !ids || !nids) return OTG_ERR_INVALID_ARG;
  otg::range_plan_result(reinterpret_cast<const otg::RangePlan *>(plan), offsets, ids, nids);
  return OTG_OK;
}

OTG_API otg_status otg_range_plan_copy(const otg_range_plan *plan, uint64_t *offsets, uint32_t *ids) {
  if (!plan || !offsets) return OTG_ERR_INVALID_ARG;
  auto *p = reinterpret_cast<const otg::RangePlan *>(plan);
  Ctx *ctx = otg::range_plan_ctx(const_cast<otg::RangePlan *>(p));
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    const uint64_t *d_off;
    const uint32_t *d_ids;
    uint64_t nids = 0;
    otg::range_plan_result(p, &d_off, &d_ids, &nids);
    if (nids && !ids) throw Error{OTG_ERR_INVALID_ARG, "ids buffer"};
    OTG_HIP_CHECK(hipMemcpyAsync(offsets, d_off, (otg::range_plan_queries(p) + 1) * 8, hipMemcpyDeviceToHost,
                                 ctx->stream));
    if (nids) OTG_HIP_CHECK(hipMemcpyAsync(ids, d_ids, nids * 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

OTG_API void otg_range_plan_free(otg_range_plan *plan) {
  if (!plan) return;
  auto *p = reinterpret_cast<otg::RangePlan *>(plan);
  Ctx *ctx = otg::range_plan_ctx(p);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  otg::range_plan_free(p);
}

OTG_API otg_status otg_step_plan_run(otg_step_plan *step) {
  if (!step) return OTG_ERR_INVALID_ARG;
  auto *st = reinterpret_cast<otg::StepPlan *>(step);
  Ctx *ctx = otg::step_plan_ctx(st);
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    otg::step_plan_run(st);
  });
}

OTG_API void otg_step_plan_free(otg_step_plan *step) {
  if (!step) return;
  auto *st = reinterpret_cast<otg::StepPlan *>(step);
  Ctx *ctx = otg::step_plan_ctx(st);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  otg::step_plan_free(st);
}

OTG_API otg_status otg_knn_plan_result(const otg_knn_plan *plan, const uint64_t **offsets, const uint32_t **ids,
                                       uint64_t *nids) {
  if (!plan || !offsets || !ids || !nids) return OTG_ERR_INVALID_ARG;
  otg::knn_plan_result(reinterpret_cast<const otg::KnnPlan *>(plan), offsets, ids, nids);
  return OTG_OK;
}

OTG_API otg_status otg_knn_plan_copy(const otg_knn_plan *plan, uint64_t *offsets, uint32_t *ids) {
  if (!plan || !offsets) return OTG_ERR_INVALID_ARG;
  auto *p = reinterpret_cast<const otg::KnnPlan *>(plan);
  Ctx *ctx = otg::knn_plan_ctx(const_cast<otg::KnnPlan *>(p));
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    const uint64_t *d_off;
    const uint32_t *d_ids;
    uint64_t nids = 0;
    otg::knn_plan_result(p, &d_off, &d_ids, &nids);
    if (nids && !ids) throw Error{OTG_ERR_INVALID_ARG, "ids buffer"};
    OTG_HIP_CHECK(hipMemcpyAsync(offsets, d_off, (otg::knn_plan_queries(p) + 1) * 8, hipMemcpyDeviceToHost,
                                 ctx->stream));
    if (nids) OTG_HIP_CHECK(hipMemcpyAsync(ids, d_ids, nids * 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

OTG_API void otg_knn_plan_free(otg_knn_plan *plan) {
  if (!plan) return;
  auto *p = reinterpret_cast<otg::KnnPlan *>(plan);
  (void)hipSetDevice(otg::knn_plan_ctx(p)->device);
  otg::knn_plan_free(p);
}

OTG_API otg_status otg_knn_footprint(const otg_tree *t, uint64_t nq, const double *points, const uint32_t *set_ids,
                                     int mem, uint32_t k, double max_within, double tolerance, uint64_t *out) {
  if (!t || !out || (nq && !points)) return OTG_ERR_INVALID_ARG;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    uint64_t fp[2] = {0, 0};
    std::unique_ptr<otg::Result> r(otg::query_knn(t, nq, points, set_ids, mem, k, max_within, tolerance, fp));
    out[0] = fp[0];
    out[1] = fp[1];
    out[2] = r->nids;
  });
}

OTG_API otg_status otg_range(const otg_tree *t, uint64_t nq, const double *ranges, const uint32_t *set_ids, int mem,
                             int mode, double tolerance, otg_result **out) {
  if (!t || !out || (nq && !ranges)) return OTG_ERR_INVALID_ARG;
  if (mode != OTG_RANGE_INSIDE && mode != OTG_RANGE_OVERLAP) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(otg::query_range(t, nq, ranges, set_ids, mem, mode, tolerance));
  });
}

OTG_API otg_status otg_ray_first(const otg_tree *t, uint64_t nq, const double *rays, const uint32_t *set_ids, int mem,
                                 double tolerance, double tolerance_increment, double max_distance, otg_result **out) {
  if (!t || !out || (nq && !rays)) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(
        otg::query_ray(t, nq, rays, set_ids, mem, tolerance, tolerance_increment, max_distance));
  });
}

OTG_API otg_status otg_collide(const otg_tree *t, uint32_t set, double tolerance, otg_result **out) {
  if (!t || !out) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(otg::query_collide(t, set, tolerance));
  });
}

OTG_API otg_status otg_collide_trees(const otg_tree *left, uint32_t left_set, const otg_tree *right,
                                     uint32_t right_set, double tolerance, otg_result **out) {
  if (!left || !right || !out) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = left->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(otg::query_collide_trees(left, left_set, right, right_set, tolerance));
  });
}

OTG_API otg_status otg_pick(const otg_tree *t, uint64_t nq, const double *points, const uint32_t *set_ids, int mem,
                            double tolerance, otg_result **out) {
  if (!t || !out || (nq && !points) || !(tolerance >= 0.0)) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(otg::query_pick(t, nq, points, set_ids, mem, tolerance));
  });
}

OTG_API otg_status otg_planes(const otg_tree *t, uint64_t nq, uint32_t nplanes, const double *planes,
                              const uint32_t *set_ids, int mem, int mode, double tolerance, otg_result **out) {
  if (!t || !out || (nq && nplanes && !planes)) return OTG_ERR_INVALID_ARG;
  if (mode != OTG_PLANE_SEARCH && mode != OTG_PLANE_POSITIVE && mode != OTG_PLANE_FRUSTUM) return OTG_ERR_INVALID_ARG;
  if (mode != OTG_PLANE_FRUSTUM && nplanes != 1) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(otg::query_planes(t, nq, nplanes, planes, set_ids, mem, mode, tolerance));
  });
}

OTG_API otg_status otg_ray_all(const otg_tree *t, uint64_t nq, const double *rays, const uint32_t *set_ids, int mem,
                               double tolerance, double tolerance_increment, double max_distance, otg_result **out) {
  if (!t || !out || (nq && !rays)) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  Ctx *ctx = t->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    *out = static_cast<otg_result *>(
        otg::query_ray_all(t, nq, rays, set_ids, mem, tolerance, tolerance_increment, max_distance));
  });
}

OTG_API otg_status otg_result_get_info(const otg_result *r, uint64_t *nq, uint64_t *nids) {
  if (!r) return OTG_ERR_INVALID_ARG;
  if (nq) *nq = r->nq;
  if (nids) *nids = r->nids;
  return OTG_OK;
}

OTG_API otg_status otg_result_copy(const otg_result *r, uint64_t *offsets, uint32_t *ids) {
  if (!r) return OTG_ERR_INVALID_ARG;
  if (r->host_only) {
    if (offsets) std::memcpy(offsets, r->h_offsets.data(), (r->nq + 1) * 8);
    if (ids && r->nids) std::memcpy(ids, r->h_ids.data(), r->nids * 4);
    return OTG_OK;
  }
  Ctx *ctx = r->ctx;
  return guarded(ctx, [&] {
    OTG_HIP_CHECK(hipSetDevice(ctx->device));
    if (offsets) OTG_HIP_CHECK(hipMemcpyAsync(offsets, r->d_offsets, (r->nq + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (ids && r->nids) OTG_HIP_CHECK(hipMemcpyAsync(ids, r->d_ids, r->nids * 4, hipMemcpyDeviceToHost, ctx->stream));
    OTG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

OTG_API otg_status otg_result_device(const otg_result *r, const uint64_t **offsets, const uint32_t **ids) {
  if (!r || r->host_only) return OTG_ERR_INVALID_ARG;
  if (offsets) *offsets = r->d_offsets;
  if (ids) *ids = r->d_ids;
  return OTG_OK;
}

OTG_API void otg_result_free(otg_result *r) {
  if (!r) return;
  delete r;
}

}  // extern "C"
