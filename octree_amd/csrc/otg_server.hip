// otg_server.hip -- serving: many resident trees under an HBM budget, continuous
// batching of query requests.
//
// The reference has no serving layer (it is a header-only library); this is the
// runtime the batched path needs on the device side:
//  * trees are registered with a host copy of their entities and built into HBM;
//    admission is by memory accounting against the budget: a tree that does not
//    fit evicts the least-recently-used trees that no running batch uses, and an
//    evicted tree is rebuilt when a request for it arrives;
//  * requests (a batch of kNN / range / ray queries on one tree) are queued and
//    taken by worker slots as they free up; a worker takes everything pending (up
//    to a query budget) and serves it with one launch per (query kind, tree kind,
//    dimension, parameters): queries on different trees share the launch (each
//    query carries its tree index; kNN queries carry their own k);
//  * two or more workers, each with its own stream and allocator, take turns, so
//    one batch is prepared and launched while the previous one is on the GPU.
// A query's ids do not depend on what it was batched with: every query is
// computed independently by the same traversal.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "otg_internal.hpp"

namespace otg {

struct ServerRequest {
  uint64_t id = 0;
  uint32_t tree_id = 0;
  int query = 0;  // 0 kNN, 1 range, 2 ray
  uint32_t k = 0;
  double p0 = 0, p1 = 0, p2 = 0;
  int mode = 0;
  uint64_t nq = 0;
  std::vector<double> data;
  double t_submit = 0, t_start = 0, t_done = 0;
  bool cold = false;            // its tree was not resident when its batch formed
  double batch_resident_ms = 0;  // wall time its batch spent making trees resident
  uint32_t batch_size = 0;      // requests in its batch
  bool done = false;
  otg_status status = OTG_OK;
  std::string error;
  otg_result *result = nullptr;
};

struct TreeFree {
  void operator()(otg_tree *t) const { otg_tree_free(t); }
};

// Host copy of a registered tree's entities in pinned memory, so a rebuild after
// eviction uploads at full link speed.
struct PinnedGeom {
  double *p = nullptr;
  uint64_t count = 0;
  PinnedGeom() = default;
  PinnedGeom(const PinnedGeom &) = delete;
  PinnedGeom &operator=(const PinnedGeom &) = delete;
  PinnedGeom(PinnedGeom &&o) noexcept : p(o.p), count(o.count) {
    o.p = nullptr;
    o.count = 0;
  }
  PinnedGeom &operator=(PinnedGeom &&o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      count = o.count;
      o.p = nullptr;
      o.count = 0;
    }
    return *this;
  }
  ~PinnedGeom() { reset(); }
  bool allocate(uint64_t n) {
    reset();
    if (n == 0) return true;
    if (hipHostMalloc((void **)&p, n * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return false;
    }
    count = n;
    return true;
  }
  void reset() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    count = 0;
  }
  double *data() { return p; }
};

// Per-worker pinned staging for query inputs and results (grows, never shrinks).
struct PinnedBytes {
  char *p = nullptr;
  uint64_t cap = 0;
  PinnedBytes() = default;
  PinnedBytes(const PinnedBytes &) = delete;
  PinnedBytes &operator=(const PinnedBytes &) = delete;
  ~PinnedBytes() {
    if (p) (void)hipHostFree(p);
  }
  char *ensure(uint64_t bytes) {
    if (bytes <= cap) return p;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const uint64_t want = std::max<uint64_t>(bytes + bytes / 4, 1ull << 20);
    if (hipHostMalloc((void **)&p, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw std::bad_alloc();
    }
    cap = want;
    return p;
  }
};

struct Staging {
  PinnedBytes in, out;
};

struct ServedTree {
  // registration (kept while evicted, to rebuild on demand)
  int kind = 0, dim = 0;
  uint64_t n = 0;
  PinnedGeom geom;
  std::vector<double> box;  // empty: bounding box of the entities
  uint32_t max_depth = 0, cap = 0;
  // residency
  std::unique_ptr<otg_tree, TreeFree> tree;
  uint64_t bytes = 0;
  uint64_t last_use = 0;
  double heat = 0.0;  // requests, decayed per batch: eviction keeps the frequently used trees
  uint32_t inflight = 0;
  bool building = false;
  uint64_t builds = 0;
};

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Admission estimate of a tree's device bytes (entities, keys, nodes and a finer
// index for point trees), used before the build reports the real figure.
static uint64_t tree_estimate(int kind, int dim, uint64_t n) {
  const uint64_t per = (kind == OTG_KIND_BOX ? 16ull * dim : 8ull * dim) + 12;
  return n * per * (kind == OTG_KIND_BOX ? 2 : 3) + (64ull << 20);
}

class Server {
 public:
  Ctx *build_ctx = nullptr;  // builds and owns tree memory
  std::vector<Ctx *> worker_ctx;
  std::vector<std::unique_ptr<Staging>> staging;  // one per worker
  uint64_t budget = 0;
  uint64_t resident = 0;  // bytes of resident trees + reservations of builds in progress
  uint64_t max_batch_queries = 1ull << 21;
  double defer_ms = 5.0;
  std::mutex mu, build_mu;
  std::condition_variable cv_work, cv_done, cv_mem;
  std::deque<ServerRequest *> pending;
  std::unordered_map<uint64_t, ServerRequest *> requests;
  std::unordered_map<uint32_t, ServedTree> trees;
  std::vector<std::thread> workers;
  bool stop = false;
  uint64_t next_id = 1, tick = 0;
  // statistics
  uint64_t batches = 0, launches = 0, served_requests = 0, served_queries = 0, evictions = 0, rebuilds = 0;
  double build_ms = 0.0, query_ms = 0.0;  // wall time in builds and in query launches (incl. copies)
  double stage_ms = 0.0, launch_ms = 0.0, split_ms = 0.0;  // query_ms split: inputs, launch, results

  void shutdown() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_work.notify_all();
    cv_done.notify_all();
    for (auto &w : workers) w.join();
    workers.clear();
    for (auto &kv : requests) {
      otg_result_free(kv.second->result);
      delete kv.second;
    }
    requests.clear();
    pending.clear();
    trees.clear();
  }

  // Evict the coldest idle trees (least requested, then least recently used)
  // until `need` more bytes fit (mu held;
  // `self` is the tree being admitted).  A caller that holds pinned trees waits a
  // bounded time (`bounded`), so two workers never wait on each other's pins.
  bool make_room(std::unique_lock<std::mutex> &lk, uint64_t need, const ServedTree *self, bool bounded) {
    if (need > budget) return false;
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(20);
    for (;;) {
      if (resident + need <= budget) return true;
      ServedTree *victim = nullptr;
      for (auto &kv : trees) {
        ServedTree &e = kv.second;
        if (e.tree && e.inflight == 0 && !e.building &&
            (!victim || e.heat < victim->heat || (e.heat == victim->heat && e.last_use < victim->last_use)))
          victim = &e;
      }
      if (victim) {
        resident -= victim->bytes;
        victim->bytes = 0;
        victim->tree.reset();
        ++evictions;
        continue;
      }
      // memory held by running batches or other builds comes back when they end
      bool busy = false;
      for (auto &kv : trees)
        if (&kv.second != self) busy |= kv.second.inflight > 0 || kv.second.building;
      if (!busy) return false;
      if (!bounded) {
        cv_mem.wait(lk);
      } else if (cv_mem.wait_until(lk, deadline) == std::cv_status::timeout) {
        return resident + need <= budget;
      }
    }
  }

  // Build e's tree into HBM (mu held on entry and exit; released while building).
  otg_status make_resident(std::unique_lock<std::mutex> &lk, ServedTree &e, bool bounded = false) {
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(20);
    for (;;) {  // another thread may be building this tree: wait for it
      if (e.tree) return OTG_OK;
      if (!e.building) break;
      if (!bounded) {
        cv_mem.wait(lk);
      } else if (cv_mem.wait_until(lk, deadline) == std::cv_status::timeout && e.building) {
        return OTG_ERR_OOM;
      }
    }
    const uint64_t est = tree_estimate(e.kind, e.dim, e.n);
    e.building = true;  // claims the build (others wait above)
    if (!make_room(lk, est, &e, bounded)) {
      e.building = false;
      cv_mem.notify_all();
      return OTG_ERR_OOM;
    }
    resident += est;
    lk.unlock();
    otg_tree *t = nullptr;
    otg_status st;
    const double tb0 = now_ms();
    {
      std::lock_guard<std::mutex> bg(build_mu);
      otg_ctx *bc = static_cast<otg_ctx *>(build_ctx);
      st = otg_tree_build(bc, e.kind, e.dim, e.n, e.geom.data(), OTG_MEM_HOST, e.box.empty() ? nullptr : e.box.data(),
                          e.max_depth, e.cap, &t);
      if (st == OTG_OK) st = otg_ctx_synchronize(bc);  // workers read the tree from their own streams
    }
    lk.lock();
    build_ms += now_ms() - tb0;
    resident -= est;
    if (st == OTG_OK) {
      otg_tree_info ti{};
      otg_tree_get_info(t, &ti);
      e.tree.reset(t);
      e.bytes = ti.device_bytes;
      resident += e.bytes;
      e.builds++;
      if (e.builds > 1) ++rebuilds;
    } else if (t) {
      otg_tree_free(t);
    }
    e.building = false;
    cv_mem.notify_all();
    return st;
  }

  void worker_loop(int w) {
    Ctx *ctx = worker_ctx[w];
    (void)hipSetDevice(ctx->device);
    set_thread_run_ctx(ctx);
    for (;;) {
      std::vector<ServerRequest *> batch;
      {
        std::unique_lock<std::mutex> lk(mu);
        // Requests for evicted trees wait up to defer_ms while other work exists, so
        // one rebuild serves every request that arrived for the tree meanwhile.
        for (;;) {
          cv_work.wait(lk, [&] { return stop || !pending.empty(); });
          if (stop) return;
          const double now = now_ms();
          uint64_t q = 0;
          double next_due = 1e300;
          for (auto it = pending.begin(); it != pending.end();) {
            ServerRequest *r = *it;
            if (!batch.empty() && q + r->nq > max_batch_queries) break;
            auto tr = trees.find(r->tree_id);
            const bool cold = tr != trees.end() && !tr->second.tree;
            if (cold && now - r->t_submit < defer_ms) {
              next_due = std::min(next_due, r->t_submit + defer_ms);
              ++it;
              continue;
            }
            q += r->nq;
            r->t_start = now;
            r->cold = cold;
            batch.push_back(r);
            it = pending.erase(it);
          }
          if (!batch.empty()) break;
          cv_work.wait_for(lk, std::chrono::duration<double, std::milli>(std::max(0.05, next_due - now)));
        }
        ++batches;
        for (auto &kv : trees) kv.second.heat *= 0.9;
      }
      const double resident_ms = run_batch(ctx, *staging[w], batch);
      {
        std::lock_guard<std::mutex> g(mu);
        for (ServerRequest *r : batch) {
          r->batch_resident_ms = resident_ms;
          r->batch_size = (uint32_t)batch.size();
          r->t_done = now_ms();
          r->done = true;
          served_requests++;
          served_queries += r->nq;
        }
      }
      cv_done.notify_all();
    }
  }

  // Serve a batch: its trees are taken in chunks -- the resident ones first, then
  // the others in chunks that fit the budget -- so each evicted tree is rebuilt at
  // most once per batch; within a chunk, one launch per (query, tree kind, dim,
  // k class, parameters).
  // Returns the wall time spent making trees resident (rebuilds and waits for memory).
  double run_batch(Ctx *ctx, Staging &stg, std::vector<ServerRequest *> &batch) {
    double resident_ms = 0.0;
    std::unordered_map<uint32_t, std::vector<ServerRequest *>> by_tree;
    std::vector<uint32_t> tree_order;
    std::vector<std::vector<uint32_t>> chunks;
    {
      std::lock_guard<std::mutex> g(mu);
      for (ServerRequest *r : batch) {
        auto it = trees.find(r->tree_id);
        if (it == trees.end()) {
          r->status = OTG_ERR_INVALID_ARG;
          r->error = "tree not registered";
          continue;
        }
        const ServedTree &e = it->second;
        if (r->query == 0 && r->k > 64) {
          r->status = OTG_ERR_UNSUPPORTED;
          r->error = "k > 64";
          continue;
        }
        auto &v = by_tree[r->tree_id];
        if (v.empty()) tree_order.push_back(r->tree_id);
        v.push_back(r);
      }
      std::vector<uint32_t> resident_now, others;
      for (uint32_t id : tree_order) (trees[id].tree ? resident_now : others).push_back(id);
      if (!resident_now.empty()) chunks.push_back(resident_now);
      std::vector<uint32_t> cur;
      uint64_t acc = 0;
      for (uint32_t id : others) {
        const ServedTree &e = trees[id];
        const uint64_t est = tree_estimate(e.kind, e.dim, e.n);
        if (!cur.empty() && acc + est > budget / 2) {
          chunks.push_back(cur);
          cur.clear();
          acc = 0;
        }
        cur.push_back(id);
        acc += est;
      }
      if (!cur.empty()) chunks.push_back(cur);
    }
    using Key = std::tuple<int, int, int, int, double, double, double, int>;
    for (const auto &chunk : chunks) {
      std::map<Key, std::vector<ServerRequest *>> groups;
      {
        std::lock_guard<std::mutex> g(mu);
        for (uint32_t id : chunk) {
          const ServedTree &e = trees[id];
          for (ServerRequest *r : by_tree[id]) {
            const int kcls = r->query == 0 ? (r->k <= 8 ? 0 : r->k <= 16 ? 1 : r->k <= 32 ? 2 : 3) : 0;
            groups[Key{r->query, e.kind, e.dim, kcls, r->p0, r->p1, r->p2, r->mode}].push_back(r);
          }
        }
      }
      for (auto &kv : groups) run_group(ctx, stg, kv.second, resident_ms);
    }
    return resident_ms;
  }

  void run_group(Ctx *ctx, Staging &stg, std::vector<ServerRequest *> &group, double &resident_ms) {
    size_t i = 0;
    while (i < group.size()) {
      // make the trees resident and pin them for the launch; when memory for the
      // next tree is held by this launch's own pins, launch what is pinned first
      std::vector<const Tree *> views;
      std::vector<uint32_t> pinned;
      std::unordered_map<uint32_t, uint32_t> view_of_tree;
      std::vector<ServerRequest *> live;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (; i < group.size(); ++i) {
          ServerRequest *r = group[i];
          if (!view_of_tree.count(r->tree_id)) {
            auto it = trees.find(r->tree_id);
            const double tr = now_ms();
            otg_status st = it == trees.end() ? OTG_ERR_INVALID_ARG : make_resident(lk, it->second, !pinned.empty());
            resident_ms += now_ms() - tr;
            it = trees.find(r->tree_id);  // the map may have changed while unlocked
            if (st != OTG_OK || it == trees.end() || !it->second.tree) {
              if (st == OTG_ERR_OOM && !pinned.empty()) break;  // retried with no pins held
              r->status = st != OTG_OK ? st : OTG_ERR_INVALID_ARG;
              r->error = "tree could not be made resident";
              continue;
            }
            ServedTree &e = it->second;
            e.inflight++;
            e.last_use = ++tick;
            pinned.push_back(r->tree_id);
            view_of_tree.emplace(r->tree_id, (uint32_t)views.size());
            views.push_back(e.tree.get());
          }
          live.push_back(r);
        }
      }
      if (!live.empty()) launch_group(ctx, stg, live, views, view_of_tree);
      {
        std::lock_guard<std::mutex> g(mu);
        for (uint32_t id : pinned) {
          auto it = trees.find(id);
          if (it != trees.end() && it->second.inflight) it->second.inflight--;
        }
      }
      cv_mem.notify_all();
    }
  }

  void launch_group(Ctx *ctx, Staging &stg, std::vector<ServerRequest *> &group,
                    const std::vector<const Tree *> &views,
                    const std::unordered_map<uint32_t, uint32_t> &view_of_tree) {
    const double t0 = now_ms();
    struct Timer {
      Server *s;
      double t0;
      ~Timer() {
        std::lock_guard<std::mutex> g(s->mu);
        s->query_ms += now_ms() - t0;
      }
    } timer{this, t0};
    const ServerRequest *a = group[0];
    const int dim = views[0]->dim;
    const int stride = a->query == 0 ? dim : 2 * dim;
    uint64_t total = 0;
    for (auto *r : group) total += r->nq;
    try {
      // inputs straight into pinned memory: queries, per-query tree index, per-query k
      const uint64_t qbytes = total * stride * 8, vbytes = total * 4;
      char *in = stg.in.ensure(qbytes + 2 * vbytes);
      double *q = reinterpret_cast<double *>(in);
      uint32_t *qv = reinterpret_cast<uint32_t *>(in + qbytes);
      uint32_t *qk = a->query == 0 ? qv + total : nullptr;
      uint64_t at = 0;
      for (auto *r : group) {
        std::copy(r->data.begin(), r->data.end(), q + at * stride);
        const uint32_t v = view_of_tree.at(r->tree_id);
        std::fill(qv + at, qv + at + r->nq, v);
        if (qk) std::fill(qk + at, qk + at + r->nq, r->k);
        at += r->nq;
      }
      const double t1 = now_ms();
      std::unique_ptr<Result> res(
          query_multi(ctx, a->query, views, total, q, qv, qk, a->p0, a->p1, a->p2, a->mode));
      const double t2 = now_ms();
      char *outb = stg.out.ensure((total + 1) * 8 + res->nids * 4);
      uint64_t *off = reinterpret_cast<uint64_t *>(outb);
      uint32_t *ids = reinterpret_cast<uint32_t *>(outb + (total + 1) * 8);
      OTG_HIP_CHECK(hipMemcpyAsync(off, res->d_offsets, (total + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
      if (res->nids) OTG_HIP_CHECK(hipMemcpyAsync(ids, res->d_ids, res->nids * 4, hipMemcpyDeviceToHost, ctx->stream));
      OTG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      {
        std::lock_guard<std::mutex> g(mu);
        launches++;
      }
      uint64_t q0 = 0;
      for (auto *r : group) {
        otg_result *out = new otg_result();
        out->host_only = true;
        out->nq = r->nq;
        const uint64_t b = off[q0], e = off[q0 + r->nq];
        out->nids = e - b;
        out->h_offsets.resize(r->nq + 1);
        for (uint64_t i = 0; i <= r->nq; ++i) out->h_offsets[i] = off[q0 + i] - b;
        out->h_ids.assign(ids + b, ids + e);
        r->result = out;
        q0 += r->nq;
      }
      {
        std::lock_guard<std::mutex> g(mu);
        stage_ms += t1 - t0;
        launch_ms += t2 - t1;
        split_ms += now_ms() - t2;
      }
    } catch (const Error &e) {
      drain(ctx);
      for (auto *r : group) {
        r->status = e.status;
        r->error = e.msg;
      }
    } catch (const std::bad_alloc &) {
      drain(ctx);
      for (auto *r : group) r->status = OTG_ERR_OOM;
    } catch (...) {
      drain(ctx);
      for (auto *r : group) r->status = OTG_ERR_HIP;
    }
  }

  // A failed launch may have left kernels queued that still read the pinned trees:
  // the worker's stream drains before the pins are released (its error, if any, is
  // the one already reported).
  static void drain(Ctx *ctx) { (void)hipStreamSynchronize(ctx->stream); }
};

}  // namespace otg

struct otg_server : public otg::Server {};

extern "C" {

OTG_API otg_status otg_server_create(int device, uint64_t hbm_budget_bytes, uint32_t nworkers, otg_server **out) {
  if (!out || nworkers == 0 || nworkers > 8) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  otg_server *s = new (std::nothrow) otg_server();
  if (!s) return OTG_ERR_OOM;
  s->budget = hbm_budget_bytes;
  otg_ctx *c = nullptr;
  otg_status st = otg_ctx_create(device, &c);
  if (st != OTG_OK) {
    delete s;
    return st;
  }
  s->build_ctx = c;
  for (uint32_t w = 0; w < nworkers; ++w) {
    otg_ctx *wc = nullptr;
    st = otg_ctx_create(device, &wc);
    if (st != OTG_OK) {
      for (auto *x : s->worker_ctx) otg_ctx_destroy(static_cast<otg_ctx *>(x));
      otg_ctx_destroy(c);
      delete s;
      return st;
    }
    s->worker_ctx.push_back(wc);
  }
  for (uint32_t w = 0; w < nworkers; ++w) s->staging.emplace_back(new otg::Staging());
  for (uint32_t w = 0; w < nworkers; ++w) s->workers.emplace_back([s, w] { s->worker_loop((int)w); });
  *out = s;
  return OTG_OK;
}

OTG_API void otg_server_destroy(otg_server *s) {
  if (!s) return;
  otg_ctx *bc = static_cast<otg_ctx *>(s->build_ctx);
  std::vector<otg::Ctx *> wcs = s->worker_ctx;
  s->shutdown();
  delete s;
  for (auto *w : wcs) otg_ctx_destroy(static_cast<otg_ctx *>(w));
  otg_ctx_destroy(bc);
}

OTG_API otg_status otg_server_add_tree(otg_server *s, uint32_t tree_id, int kind, int dim, uint64_t n, const double *geom,
                                       int mem, const double *box_or_null, uint32_t max_depth, uint32_t max_elements) {
  if (!s || (n && !geom) || (kind != OTG_KIND_POINT && kind != OTG_KIND_BOX) || dim < 2 || dim > otg::kMaxDim ||
      (mem != OTG_MEM_HOST && mem != OTG_MEM_DEVICE))
    return OTG_ERR_INVALID_ARG;
  otg::ServedTree e;
  e.kind = kind;
  e.dim = dim;
  e.n = n;
  e.max_depth = max_depth;
  e.cap = max_elements;
  const uint64_t cols = kind == OTG_KIND_BOX ? 2ull * dim : (uint64_t)dim;
  (void)hipSetDevice(s->build_ctx->device);
  try {
    if (box_or_null) e.box.assign(box_or_null, box_or_null + 2 * dim);
  } catch (...) {
    return OTG_ERR_OOM;
  }
  if (!e.geom.allocate(n * cols)) return OTG_ERR_OOM;
  if (n) {
    if (mem == OTG_MEM_HOST) {
      std::copy(geom, geom + n * cols, e.geom.data());
    } else if (hipMemcpy(e.geom.data(), geom, n * cols * 8, hipMemcpyDeviceToHost) != hipSuccess) {
      return OTG_ERR_HIP;
    }
  }
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->trees.count(tree_id)) {  // replace: wait until nothing uses the old one
    s->cv_mem.wait(lk, [&] {
      auto i2 = s->trees.find(tree_id);
      return i2 == s->trees.end() || (i2->second.inflight == 0 && !i2->second.building);
    });
    auto it = s->trees.find(tree_id);
    if (it != s->trees.end()) {
      s->resident -= it->second.bytes;
      s->trees.erase(it);
    }
  }
  otg::ServedTree &slot = s->trees[tree_id];
  slot = std::move(e);
  otg_status st = s->make_resident(lk, slot);
  if (st != OTG_OK) {
    auto i3 = s->trees.find(tree_id);
    if (i3 != s->trees.end() && !i3->second.tree && !i3->second.building && i3->second.inflight == 0)
      s->trees.erase(i3);
  }
  return st;
}

OTG_API otg_status otg_server_remove_tree(otg_server *s, uint32_t tree_id) {
  if (!s) return OTG_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->trees.count(tree_id)) return OTG_ERR_INVALID_ARG;
  s->cv_mem.wait(lk, [&] {
    auto i2 = s->trees.find(tree_id);
    return i2 == s->trees.end() || (i2->second.inflight == 0 && !i2->second.building);
  });
  auto it = s->trees.find(tree_id);
  if (it != s->trees.end()) {
    s->resident -= it->second.bytes;
    s->trees.erase(it);
  }
  s->cv_mem.notify_all();
  return OTG_OK;
}

OTG_API otg_status otg_server_has_tree(otg_server *s, uint32_t tree_id, int *present) {
  if (!s || !present) return OTG_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->trees.find(tree_id);
  *present = it == s->trees.end() ? 0 : (it->second.tree ? 1 : 2);
  return OTG_OK;
}

OTG_API otg_status otg_server_submit(otg_server *s, uint32_t tree_id, int query, uint64_t nq, const double *queries,
                                     uint32_t k, double p0, double p1, double p2, int mode, uint64_t *request_id) {
  if (!s || !request_id || (nq && !queries) || query < 0 || query > 2) return OTG_ERR_INVALID_ARG;
  int dim = 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->trees.find(tree_id);
    if (it == s->trees.end()) return OTG_ERR_INVALID_ARG;
    dim = it->second.dim;
    it->second.last_use = ++s->tick;  // recency and heat count requests, not only launches
    it->second.heat += 1.0;
  }
  auto *r = new (std::nothrow) otg::ServerRequest();
  if (!r) return OTG_ERR_OOM;
  r->tree_id = tree_id;
  r->query = query;
  r->k = k;
  r->p0 = p0;
  r->p1 = p1;
  r->p2 = p2;
  r->mode = mode;
  r->nq = nq;
  const uint64_t stride = query == 0 ? dim : 2 * dim;
  try {
    r->data.assign(queries, queries + nq * stride);
  } catch (...) {
    delete r;
    return OTG_ERR_OOM;
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    r->id = s->next_id++;
    r->t_submit = otg::now_ms();
    s->requests[r->id] = r;
    s->pending.push_back(r);
    *request_id = r->id;
  }
  s->cv_work.notify_one();
  return OTG_OK;
}

OTG_API otg_status otg_server_wait(otg_server *s, uint64_t request_id, otg_result **out, double *latency_ms,
                                   double *queue_ms) {
  if (!s || !out) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(request_id);
  if (it == s->requests.end()) return OTG_ERR_INVALID_ARG;
  otg::ServerRequest *r = it->second;
  s->cv_done.wait(lk, [&] { return r->done || s->stop; });
  if (!r->done) return OTG_ERR_INVALID_ARG;
  s->requests.erase(r->id);
  lk.unlock();
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  if (queue_ms) *queue_ms = r->t_start - r->t_submit;
  otg_status st = r->status;
  *out = r->result;
  r->result = nullptr;
  delete r;
  return st;
}

OTG_API otg_status otg_server_wait_detail(otg_server *s, uint64_t request_id, otg_result **out, double *detail) {
  if (!s || !out || !detail) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(request_id);
  if (it == s->requests.end()) return OTG_ERR_INVALID_ARG;
  otg::ServerRequest *r = it->second;
  s->cv_done.wait(lk, [&] { return r->done || s->stop; });
  if (!r->done) return OTG_ERR_INVALID_ARG;
  s->requests.erase(r->id);
  lk.unlock();
  detail[0] = r->t_done - r->t_submit;
  detail[1] = r->t_start - r->t_submit;
  detail[2] = r->t_done - r->t_start;
  detail[3] = r->batch_resident_ms;
  detail[4] = (double)r->batch_size;
  detail[5] = r->cold ? 1.0 : 0.0;
  otg_status st = r->status;
  *out = r->result;
  r->result = nullptr;
  delete r;
  return st;
}

OTG_API otg_status otg_server_set_batch_limit(otg_server *s, uint64_t max_queries) {
  if (!s || max_queries == 0) return OTG_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  s->max_batch_queries = max_queries;
  return OTG_OK;
}

OTG_API otg_status otg_server_stats(otg_server *s, uint64_t *stats) {
  if (!s || !stats) return OTG_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  uint64_t res_trees = 0;
  for (auto &kv : s->trees) res_trees += kv.second.tree ? 1 : 0;
  stats[0] = s->batches;
  stats[1] = s->launches;
  stats[2] = s->served_requests;
  stats[3] = s->served_queries;
  stats[4] = s->evictions;
  stats[5] = s->resident;
  stats[6] = res_trees;
  stats[7] = s->pending.size();
  stats[8] = s->rebuilds;
  stats[9] = s->trees.size();
  stats[10] = (uint64_t)(s->build_ms * 1000.0);
  stats[11] = (uint64_t)(s->query_ms * 1000.0);
  stats[12] = (uint64_t)(s->stage_ms * 1000.0);
  stats[13] = (uint64_t)(s->launch_ms * 1000.0);
  stats[14] = (uint64_t)(s->split_ms * 1000.0);
  return OTG_OK;
}

}  // extern "C"
