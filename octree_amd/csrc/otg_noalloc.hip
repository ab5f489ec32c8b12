// otg_noalloc.hip -- the allocation-free kNN entry point.
//
// For callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc): every temporary of the
// kNN batch lives in one file-scope device array, `scratch`, of fixed size, held
// by each GPU the library runs on until the process ends.  Host code never uses
// the array's name as an address: its device address is looked up once per
// device with hipGetSymbolAddress (at context creation, with the device current)
// and that pointer is what the kernels receive.
#include <mutex>

#include "otg_internal.hpp"

namespace otg {

constexpr size_t kNoAllocScratchBytes = 96ull << 20;  // 96 MiB per device
constexpr uint64_t kNoAllocMaxQueries = 1ull << 18;   // 262,144 queries per call
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_init_mu;
void *g_scratch[kNoAllocMaxDevices] = {};
hipError_t g_scratch_err[kNoAllocMaxDevices] = {};
// one call at a time per device: calls share that device's scratch array
std::mutex g_call_mu[kNoAllocMaxDevices];
}  // namespace

void noalloc_init_device(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return;
  std::lock_guard<std::mutex> g(g_init_mu);
  if (g_scratch[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return;
  }
  g_scratch[device] = p;
  g_scratch_err[device] = hipSuccess;
}

}  // namespace otg

extern "C" {

OTG_API otg_status otg_knn_noalloc(const otg_tree *t, uint64_t nq, const double *points, const uint32_t *set_ids,
                                   uint32_t k, double max_within, double tolerance, uint64_t *offsets, uint32_t *ids,
                                   uint64_t ids_capacity, uint64_t *nids) {
  using namespace otg;
  if (!t || !offsets || !nids || (nq && !points) || (ids_capacity && !ids)) return OTG_ERR_INVALID_ARG;
  *nids = 0;
  Ctx *ctx = t->ctx;
  if (nq > kNoAllocMaxQueries) {
    ctx->last_error = "otg_knn_noalloc takes at most 262,144 queries per call";
    return OTG_ERR_INVALID_ARG;
  }
  const int dev = ctx->device;
  if (dev < 0 || dev >= kNoAllocMaxDevices || !g_scratch[dev]) {
    ctx->last_error = "the scratch array's device address is not available";
    return OTG_ERR_HIP;
  }
  std::lock_guard<std::mutex> g(g_call_mu[dev]);
  // no kernel timing here: its events would be created on demand
  struct NoProfiling {
    Ctx *c;
    bool was;
    ~NoProfiling() { c->profiling = was; }
  } np{ctx, ctx->profiling};
  ctx->profiling = false;
  try {
    OTG_HIP_CHECK(hipSetDevice(dev));
    Arena ar;
    ar.base = static_cast<char *>(g_scratch[dev]);
    ar.cap = kNoAllocScratchBytes;
    CsrOut co;
    co.offsets = offsets;
    co.ids = ids;
    co.capacity = ids_capacity;
    try {
      query_knn_noalloc(t, nq, points, set_ids, k, max_within, tolerance, ar, co);
    } catch (...) {
      *nids = co.nids;  // the size needed when the capacity was short
      (void)hipStreamSynchronize(ctx->stream);
      throw;
    }
    *nids = co.nids;
    ctx->last_error.clear();
    return OTG_OK;
  } catch (const Error &e) {
    ctx->last_error = e.msg;
    return e.status;
  } catch (const std::bad_alloc &) {
    ctx->last_error = "host allocation failed";
    return OTG_ERR_OOM;
  } catch (...) {
    ctx->last_error = "unknown error";
    return OTG_ERR_HIP;
  }
}

}  // extern "C"
