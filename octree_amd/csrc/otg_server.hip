// otg_server.hip -- serving: many resident trees under an HBM budget, continuous
// batching of query requests.
//
// The reference has no serving layer (it is a header-only library); this is the
// MI355X-side runtime the batched path needs:
//  * trees are built into HBM and stay resident; adding a tree that does not
//    fit evicts least-recently-used trees no request is using, and waits for
//    in-flight work to release memory before giving up;
//  * requests (a batch of kNN / range / ray queries on one tree) are queued
//    and admitted by worker slots as they free up; a worker takes everything
//    pending (up to a query budget), coalesces requests on the same tree with
//    the same query parameters into one launch, and splits the CSR result back
//    per request;
//  * two or more workers, each with its own context (stream + allocator),
//    take turns, so a batch is built and launched while the previous one is
//    still on the GPU.
// A request's ids do not depend on what it was batched with: each query's
// result is computed independently by the same kernels.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "otg_internal.hpp"

namespace otg {

struct ServerRequest {
  uint64_t id = 0;
  uint32_t tree_id = 0;
  int query = 0;  // 0 kNN, 1 range, 2 ray
  uint32_t k = 0;
  double p0 = 0, p1 = 0, p2 = 0;  // kNN: max_within, tol ; range: tol ; ray: tol, inc, max_dist
  int mode = 0;
  uint64_t nq = 0;
  std::vector<double> data;
  double t_submit = 0, t_start = 0, t_done = 0;
  bool done = false;
  otg_status status = OTG_OK;
  std::string error;
  otg_result *result = nullptr;
};

struct TreeFree {
  void operator()(otg_tree *t) const { otg_tree_free(t); }
};

struct ServedTree {
  std::unique_ptr<otg_tree, TreeFree> tree;
  uint64_t bytes = 0;
  uint64_t last_use = 0;
  uint32_t inflight = 0;
  bool evict_pending = false;
};

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

class Server {
 public:
  Ctx *build_ctx = nullptr;  // builds and owns tree memory
  std::vector<Ctx *> worker_ctx;
  uint64_t budget = 0;
  uint64_t resident = 0;
  uint64_t max_batch_queries = 1ull << 21;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_mem;
  std::deque<ServerRequest *> pending;
  std::unordered_map<uint64_t, ServerRequest *> requests;
  std::unordered_map<uint32_t, ServedTree> trees;
  std::vector<std::thread> workers;
  bool stop = false;
  uint64_t next_id = 1, tick = 0;
  // statistics
  uint64_t batches = 0, launches = 0, served_requests = 0, served_queries = 0, evictions = 0;

  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_work.notify_all();
    for (auto &w : workers) w.join();
    for (auto &kv : requests) {
      delete kv.second->result;
      delete kv.second;
    }
    trees.clear();
  }

  void worker_loop(int w) {
    Ctx *ctx = worker_ctx[w];
    (void)hipSetDevice(ctx->device);
    set_thread_run_ctx(ctx);
    for (;;) {
      std::vector<ServerRequest *> batch;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        if (stop && pending.empty()) return;
        uint64_t q = 0;
        while (!pending.empty() && (batch.empty() || q + pending.front()->nq <= max_batch_queries)) {
          ServerRequest *r = pending.front();
          pending.pop_front();
          q += r->nq;
          r->t_start = now_ms();
          batch.push_back(r);
          auto it = trees.find(r->tree_id);
          if (it != trees.end()) {
            it->second.inflight++;
            it->second.last_use = ++tick;
          }
        }
        ++batches;
      }
      run_batch(batch);
      {
        std::lock_guard<std::mutex> g(mu);
        for (ServerRequest *r : batch) {
          auto it = trees.find(r->tree_id);
          if (it != trees.end() && it->second.inflight) it->second.inflight--;
          r->t_done = now_ms();
          r->done = true;
          served_requests++;
          served_queries += r->nq;
        }
      }
      cv_done.notify_all();
      cv_mem.notify_all();
    }
  }

  // Coalesce requests with the same (tree, query, parameters) into one launch.
  void run_batch(std::vector<ServerRequest *> &batch) {
    std::vector<bool> used(batch.size(), false);
    for (size_t i = 0; i < batch.size(); ++i) {
      if (used[i]) continue;
      ServerRequest *a = batch[i];
      std::vector<ServerRequest *> group;
      for (size_t j = i; j < batch.size(); ++j) {
        ServerRequest *b = batch[j];
        if (used[j]) continue;
        if (b->tree_id == a->tree_id && b->query == a->query && b->k == a->k && b->p0 == a->p0 && b->p1 == a->p1 &&
            b->p2 == a->p2 && b->mode == a->mode) {
          group.push_back(b);
          used[j] = true;
        }
      }
      run_group(group);
    }
  }

  void run_group(std::vector<ServerRequest *> &group) {
    Tree *t = nullptr;
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = trees.find(group[0]->tree_id);
      if (it != trees.end()) t = it->second.tree.get();
    }
    if (!t) {
      for (auto *r : group) {
        r->status = OTG_ERR_INVALID_ARG;
        r->error = "tree not resident";
      }
      return;
    }
    const int stride = (group[0]->query == 0) ? t->dim : 2 * t->dim;
    uint64_t total = 0;
    for (auto *r : group) total += r->nq;
    std::vector<double> cat;
    const double *data = nullptr;
    if (group.size() == 1) {
      data = group[0]->data.data();
    } else {
      cat.reserve(total * stride);
      for (auto *r : group) cat.insert(cat.end(), r->data.begin(), r->data.end());
      data = cat.data();
    }
    try {
      ServerRequest *a = group[0];
      std::unique_ptr<Result> res;
      if (a->query == 0) res.reset(query_knn(t, total, data, nullptr, OTG_MEM_HOST, a->k, a->p0, a->p1));
      else if (a->query == 1) res.reset(query_range(t, total, data, nullptr, OTG_MEM_HOST, a->mode, a->p0));
      else res.reset(query_ray(t, total, data, nullptr, OTG_MEM_HOST, a->p0, a->p1, a->p2));
      launches++;
      Ctx *ctx = res->ctx;
      std::vector<uint64_t> off(total + 1);
      std::vector<uint32_t> ids(res->nids);
      OTG_HIP_CHECK(hipMemcpyAsync(off.data(), res->d_offsets, (total + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
      if (res->nids)
        OTG_HIP_CHECK(hipMemcpyAsync(ids.data(), res->d_ids, res->nids * 4, hipMemcpyDeviceToHost, ctx->stream));
      OTG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      uint64_t q0 = 0;
      for (auto *r : group) {
        otg_result *out = new otg_result();
        out->host_only = true;
        out->nq = r->nq;
        const uint64_t b = off[q0], e = off[q0 + r->nq];
        out->nids = e - b;
        out->h_offsets.resize(r->nq + 1);
        for (uint64_t i = 0; i <= r->nq; ++i) out->h_offsets[i] = off[q0 + i] - b;
        out->h_ids.assign(ids.begin() + b, ids.begin() + e);
        r->result = out;
        q0 += r->nq;
      }
    } catch (const Error &e) {
      for (auto *r : group) {
        r->status = e.status;
        r->error = e.msg;
      }
    } catch (...) {
      for (auto *r : group) r->status = OTG_ERR_HIP;
    }
  }

  // Evict least-recently-used idle trees until `need` more bytes fit.
  bool make_room(std::unique_lock<std::mutex> &lk, uint64_t need) {
    if (need > budget) return false;
    for (;;) {
      if (resident + need <= budget) return true;
      uint32_t victim = 0;
      uint64_t best = UINT64_MAX;
      bool found = false;
      for (auto &kv : trees)
        if (kv.second.inflight == 0 && kv.second.last_use < best) {
          best = kv.second.last_use;
          victim = kv.first;
          found = true;
        }
      if (found) {
        resident -= trees[victim].bytes;
        trees.erase(victim);
        ++evictions;
        continue;
      }
      if (trees.empty()) return resident + need <= budget;
      cv_mem.wait(lk);  // everything resident is in use: wait for a batch to finish
    }
  }
};

}  // namespace otg

struct otg_server : public otg::Server {};

using otg::Error;

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
  for (uint32_t w = 0; w < nworkers; ++w) s->workers.emplace_back([s, w] { s->worker_loop((int)w); });
  *out = s;
  return OTG_OK;
}

OTG_API void otg_server_destroy(otg_server *s) {
  if (!s) return;
  otg_ctx *bc = static_cast<otg_ctx *>(s->build_ctx);
  std::vector<otg::Ctx *> wcs = s->worker_ctx;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &w : s->workers) w.join();
  s->workers.clear();
  for (auto &kv : s->requests) {
    delete kv.second->result;
    delete kv.second;
  }
  s->requests.clear();
  s->trees.clear();
  delete s;
  for (auto *w : wcs) otg_ctx_destroy(static_cast<otg_ctx *>(w));
  otg_ctx_destroy(bc);
}

OTG_API otg_status otg_server_add_tree(otg_server *s, uint32_t tree_id, int kind, int dim, uint64_t n, const double *geom,
                                       int mem, const double *box_or_null, uint32_t max_depth, uint32_t max_elements) {
  if (!s) return OTG_ERR_INVALID_ARG;
  otg_ctx *bc = static_cast<otg_ctx *>(s->build_ctx);
  // size estimate for admission: entities + nodes (~n/2 nodes worst case) + a finer kNN index for points
  const uint64_t per = (kind == OTG_KIND_BOX ? 16ull * dim : 8ull * dim) + 12;
  const uint64_t est = n * per * (kind == OTG_KIND_BOX ? 2 : 3) + (64ull << 20);
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->trees.find(tree_id);
    if (it != s->trees.end()) {
      if (it->second.inflight) return OTG_ERR_INVALID_ARG;
      s->resident -= it->second.bytes;
      s->trees.erase(it);
    }
    if (!s->make_room(lk, est)) return OTG_ERR_OOM;
    s->resident += est;  // reserved while building
  }
  otg_tree *t = nullptr;
  otg_status st = otg_tree_build(bc, kind, dim, n, geom, mem, box_or_null, max_depth, max_elements, &t);
  if (st == OTG_OK) st = otg_ctx_synchronize(bc);  // workers read the tree from their own streams
  if (st != OTG_OK && t) {
    otg_tree_free(t);
    t = nullptr;
  }
  std::lock_guard<std::mutex> g(s->mu);
  s->resident -= est;
  if (st != OTG_OK) return st;
  otg_tree_info ti{};
  otg_tree_get_info(t, &ti);
  otg::ServedTree e;
  e.tree.reset(t);
  e.bytes = ti.device_bytes;
  e.last_use = ++s->tick;
  s->resident += e.bytes;
  s->trees[tree_id] = std::move(e);
  return OTG_OK;
}

OTG_API otg_status otg_server_remove_tree(otg_server *s, uint32_t tree_id) {
  if (!s) return OTG_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->trees.find(tree_id);
  if (it == s->trees.end()) return OTG_ERR_INVALID_ARG;
  s->cv_mem.wait(lk, [&] { return s->trees.count(tree_id) == 0 || s->trees[tree_id].inflight == 0; });
  it = s->trees.find(tree_id);
  if (it != s->trees.end()) {
    s->resident -= it->second.bytes;
    s->trees.erase(it);
  }
  return OTG_OK;
}

OTG_API otg_status otg_server_has_tree(otg_server *s, uint32_t tree_id, int *present) {
  if (!s || !present) return OTG_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  *present = s->trees.count(tree_id) ? 1 : 0;
  return OTG_OK;
}

// query: 0 kNN (k, p0 = max_within, p1 = tolerance), 1 range (mode, p0 = tolerance),
// 2 ray first (p0 = tolerance, p1 = tolerance increment, p2 = max distance).
OTG_API otg_status otg_server_submit(otg_server *s, uint32_t tree_id, int query, uint64_t nq, const double *queries,
                                     uint32_t k, double p0, double p1, double p2, int mode, uint64_t *request_id) {
  if (!s || !request_id || (nq && !queries) || query < 0 || query > 2) return OTG_ERR_INVALID_ARG;
  int dim = 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->trees.find(tree_id);
    if (it == s->trees.end()) return OTG_ERR_INVALID_ARG;
    dim = it->second.tree->dim;
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

// Blocks until the request is done; the caller owns *out (free with otg_result_free).
OTG_API otg_status otg_server_wait(otg_server *s, uint64_t request_id, otg_result **out, double *latency_ms,
                                   double *queue_ms) {
  if (!s || !out) return OTG_ERR_INVALID_ARG;
  *out = nullptr;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(request_id);
  if (it == s->requests.end()) return OTG_ERR_INVALID_ARG;
  otg::ServerRequest *r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  s->requests.erase(it);
  lk.unlock();
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  if (queue_ms) *queue_ms = r->t_start - r->t_submit;
  otg_status st = r->status;
  *out = r->result;
  r->result = nullptr;
  delete r;
  return st;
}

OTG_API otg_status otg_server_stats(otg_server *s, uint64_t *stats /* 8 values */) {
  if (!s || !stats) return OTG_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  stats[0] = s->batches;
  stats[1] = s->launches;
  stats[2] = s->served_requests;
  stats[3] = s->served_queries;
  stats[4] = s->evictions;
  stats[5] = s->resident;
  stats[6] = s->trees.size();
  stats[7] = s->pending.size();
  return OTG_OK;
}

}  // extern "C"
