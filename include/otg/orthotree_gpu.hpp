// orthotree_gpu.hpp -- C++20 drop-in for OrthoTree's static linear trees, backed
// by libotg (MI355X / gfx950) through the C ABI in otg.h.
//
// Same class names, constructor and method signatures as the reference's
// static-tree aliases and query layer:
//   StaticOctreePoint / StaticOctreeBox / StaticQuadtree* / StaticOrthoTreePointND<D> /
//   StaticOrthoTreeBoxND<D, true>                         (core/ot_aliases.h:82-91, :225-245)
//   ctor(PAR_EXEC, entities, maxDepth, boxSpace, maxElementNo) and Create(...)
//                                                         (core/ot_static_linear_core.h:103-129, :523-613)
//   GetNearestNeighbors, RangeSearch, RayIntersectedFirst, CollisionDetection
//                                                         (core/ot_query.h:1208-1270, :476-583, :2035-2146, :1901-1906)
//   PickSearch, RayIntersectedAll, two-tree CollisionDetection
//                                                         (core/ot_query.h:426-466, :1929-2022, :1462-1496)
//   PlaneSearch / PlaneIntersection, PlanePositiveSegmentation, FrustumCulling
//                                                         (core/ot_query.h:612-822)
//   OrthoTreeManaged (entity-owning) + StaticOctreePointM / StaticOctreeBoxM / ...
//                                                         (core/ot_managed.h:47-893, core/ot_aliases.h:417-447)
//   GetNodeCount / GetNodeChildren / GetNodeEntities / GetNodeMinPoint / GetNodeSize / GetNodeBox
//                                                         (core/ot_static_linear_core.h:178-268)
//   Serialize() / Deserialize(): the BinaryOutputArchive bytes of the core / managed
//                                   tree, and loading them back (serialization.h:373-386, :433-439)
// plus batch overloads over spans of queries returning CSR results.
//
// The tree copies the entities into HBM at Create and keeps them resident; later
// calls take the entity span only for signature parity.  Entity testers and
// distance callbacks are not accepted (a host callback cannot run on the GPU).
// Errors: like the reference (noexcept, std::terminate on allocation failure),
// a failed device call prints the libotg status and terminates; Create returns
// false for points outside an explicit box (where the reference never returns).
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <optional>
#include <span>
#include <type_traits>
#include <utility>
#include <vector>

#include "otg.h"

namespace otg {

using depth_t = uint32_t;
using index_t = uint32_t;

namespace ExecutionTags {
struct Sequential {};
struct Parallel {};
}  // namespace ExecutionTags
using SeqExec = ExecutionTags::Sequential;
using ParExec = ExecutionTags::Parallel;
inline constexpr SeqExec SEQ_EXEC{};
inline constexpr ParExec PAR_EXEC{};
inline constexpr ParExec GPU_EXEC{};

template <int D>
using PointND = std::array<double, D>;
template <int D>
struct BoundingBoxND {
  PointND<D> Min, Max;
};
template <int D>
struct RayND {
  PointND<D> Origin, Direction;
};
// Plane equation dot(Normal, x) = OrigoDistance, Normal of unit length (adapters/general.h:364-368)
template <int D>
struct PlaneND {
  double OrigoDistance;
  PointND<D> Normal;
};
static_assert(sizeof(PlaneND<3>) == 4 * sizeof(double), "plane layout: origo distance, then the normal");
using Point2D = PointND<2>;
using Point3D = PointND<3>;
using BoundingBox2D = BoundingBoxND<2>;
using BoundingBox3D = BoundingBoxND<3>;

enum class RangeSearchMode { Inside, Overlap };

inline constexpr double BASE_TOLERANCE = std::numeric_limits<double>::epsilon() * 100.0;

namespace detail {
[[noreturn]] inline void fail(otg_status st, const char *what, const otg_ctx *ctx) {
  std::fprintf(stderr, "libotg: %s failed: %s (%s)\n", what, otg_status_string(st), ctx ? otg_ctx_last_error(ctx) : "");
  std::terminate();
}
inline void check(otg_status st, const char *what, const otg_ctx *ctx) {
  if (st != OTG_OK) fail(st, what, ctx);
}
}  // namespace detail

// One device + stream.  Context::Default() is shared by trees built without an explicit context.
class Context {
 public:
  explicit Context(int device = 0) { detail::check(otg_ctx_create(device, &ctx_), "otg_ctx_create", nullptr); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  ~Context() { otg_ctx_destroy(ctx_); }
  otg_ctx *get() const noexcept { return ctx_; }
  static Context &Default() {
    static Context c(0);
    return c;
  }

 private:
  otg_ctx *ctx_ = nullptr;
};

// CSR batch result: ids of query i are ids[offsets[i] .. offsets[i+1]).
struct BatchResult {
  std::vector<uint64_t> offsets;
  std::vector<index_t> ids;
  std::span<const index_t> operator[](size_t i) const {
    return std::span<const index_t>(ids.data() + offsets[i], offsets[i + 1] - offsets[i]);
  }
  size_t size() const { return offsets.empty() ? 0 : offsets.size() - 1; }
};

template <int DIM, bool IS_BOX>
class StaticOrthoTreeGPU {
 public:
  static constexpr int DIMENSION_NO = DIM;
  using TVector = PointND<DIM>;
  using TBox = BoundingBoxND<DIM>;
  using TRay = RayND<DIM>;
  using TPlane = PlaneND<DIM>;
  using Entity = std::conditional_t<IS_BOX, TBox, TVector>;
  using EntityContainerView = std::span<const Entity>;
  using NodeID = uint32_t;
  using EntityID = index_t;

  StaticOrthoTreeGPU() = default;
  explicit StaticOrthoTreeGPU(Context &ctx) noexcept : ctx_(&ctx) {}

  template <typename TExecMode>
  StaticOrthoTreeGPU(TExecMode execMode, EntityContainerView entities, std::optional<depth_t> maxDepthID = std::nullopt,
                     std::optional<TBox> boxSpace = std::nullopt, std::size_t maxElementNoInNode = 16) noexcept {
    [[maybe_unused]] bool ok = Create(entities, maxDepthID, boxSpace, maxElementNoInNode, execMode);
  }
  // without an execution tag the reference builds with SeqExec (ot_static_linear_core.h:105-115)
  explicit StaticOrthoTreeGPU(EntityContainerView entities, std::optional<depth_t> maxDepthID = std::nullopt,
                              std::optional<TBox> boxSpace = std::nullopt, std::size_t maxElementNoInNode = 16) noexcept {
    [[maybe_unused]] bool ok = Create(entities, maxDepthID, boxSpace, maxElementNoInNode, SEQ_EXEC);
  }
  StaticOrthoTreeGPU(const StaticOrthoTreeGPU &) = delete;
  StaticOrthoTreeGPU &operator=(const StaticOrthoTreeGPU &) = delete;
  StaticOrthoTreeGPU(StaticOrthoTreeGPU &&o) noexcept { *this = std::move(o); }
  StaticOrthoTreeGPU &operator=(StaticOrthoTreeGPU &&o) noexcept {
    std::swap(ctx_, o.ctx_);
    std::swap(tree_, o.tree_);
    std::swap(nodes_, o.nodes_);
    std::swap(entity_count_, o.entity_count_);
    return *this;
  }
  ~StaticOrthoTreeGPU() { otg_tree_free(tree_); }

  // ---- build ------------------------------------------------------------------
  // Create<ParExec> numbers the tree by sorted locations, Create<SeqExec> (the
  // reference's default) as the unsorted partition build does (:417-519).
  template <typename TExecMode = SeqExec>
  bool Create(EntityContainerView entities, std::optional<depth_t> maxDepthIn = std::nullopt,
              std::optional<TBox> boxSpace = std::nullopt, std::size_t maxElementNoInNode = 16,
              TExecMode = {}) noexcept {
    otg_tree_free(tree_);
    tree_ = nullptr;
    nodes_ = {};
    std::array<double, 2 * DIM> box{};
    if (boxSpace)
      for (int d = 0; d < DIM; ++d) {
        box[d] = boxSpace->Min[d];
        box[DIM + d] = boxSpace->Max[d];
      }
    auto build = std::is_same_v<TExecMode, SeqExec> ? otg_tree_build_seq : otg_tree_build;
    otg_status st = build(ctx().get(), IS_BOX ? OTG_KIND_BOX : OTG_KIND_POINT, DIM, entities.size(),
                          reinterpret_cast<const double *>(entities.data()), OTG_MEM_HOST,
                          boxSpace ? box.data() : nullptr, maxDepthIn.value_or(0),
                          static_cast<uint32_t>(maxElementNoInNode), &tree_);
    if (st == OTG_ERR_OUT_OF_SPACE) return false;
    detail::check(st, "otg_tree_build", ctx().get());
    entity_count_ = entities.size();
    return true;
  }

  template <typename TExecMode = SeqExec>
  static bool Create(StaticOrthoTreeGPU &tree, EntityContainerView entities, std::optional<depth_t> maxDepthIn = std::nullopt,
                     std::optional<TBox> boxSpace = std::nullopt, std::size_t maxElementNoInNode = 16,
                     TExecMode execMode = {}) noexcept {
    return tree.Create(entities, maxDepthIn, boxSpace, maxElementNoInNode, execMode);
  }

  // ---- queries ------------------------------------------------------------------
  std::vector<EntityID> GetNearestNeighbors(TVector const &searchPoint, std::size_t neighborNo, EntityContainerView entities,
                                            double tolerance = BASE_TOLERANCE) const noexcept
  {
    return GetNearestNeighbors(searchPoint, neighborNo, std::numeric_limits<double>::max(), entities, tolerance);
  }
  std::vector<EntityID> GetNearestNeighbors(TVector const &searchPoint, std::size_t neighborNo, double maxDistanceWithin,
                                            EntityContainerView, double tolerance = BASE_TOLERANCE) const noexcept
  {
    auto r = GetNearestNeighbors(std::span<const TVector>(&searchPoint, 1), neighborNo, maxDistanceWithin, tolerance);
    return std::vector<EntityID>(r.ids.begin(), r.ids.end());
  }
  // batch overload
  BatchResult GetNearestNeighbors(std::span<const TVector> searchPoints, std::size_t neighborNo,
                                  double maxDistanceWithin = std::numeric_limits<double>::max(),
                                  double tolerance = BASE_TOLERANCE) const noexcept
  {
    otg_result *res = nullptr;
    if (!tree_) return empty(searchPoints.size());
    detail::check(otg_knn(tree_, searchPoints.size(), reinterpret_cast<const double *>(searchPoints.data()), nullptr,
                          OTG_MEM_HOST, static_cast<uint32_t>(neighborNo), maxDistanceWithin, tolerance, &res),
                  "otg_knn", ctx().get());
    return take(res);
  }

  std::vector<EntityID> RangeSearch(TBox const &range, EntityContainerView, RangeSearchMode mode = RangeSearchMode::Inside,
                                    double tolerance = BASE_TOLERANCE) const noexcept {
    auto r = RangeSearch(std::span<const TBox>(&range, 1), mode, tolerance);
    return std::vector<EntityID>(r.ids.begin(), r.ids.end());
  }
  BatchResult RangeSearch(std::span<const TBox> ranges, RangeSearchMode mode = RangeSearchMode::Inside,
                          double tolerance = BASE_TOLERANCE) const noexcept {
    otg_result *res = nullptr;
    if (!tree_) return empty(ranges.size());
    detail::check(otg_range(tree_, ranges.size(), reinterpret_cast<const double *>(ranges.data()), nullptr, OTG_MEM_HOST,
                            mode == RangeSearchMode::Inside ? OTG_RANGE_INSIDE : OTG_RANGE_OVERLAP, tolerance, &res),
                  "otg_range", ctx().get());
    return take(res);
  }

  std::vector<EntityID> RayIntersectedFirst(TVector const &rayBasePoint, TVector const &rayHeading, EntityContainerView,
                                            double tolerance = BASE_TOLERANCE, double toleranceIncrement = 0.0,
                                            double maxDistance = std::numeric_limits<double>::max()) const noexcept {
    TRay ray{rayBasePoint, rayHeading};
    auto r = RayIntersectedFirst(std::span<const TRay>(&ray, 1), tolerance, toleranceIncrement, maxDistance);
    return std::vector<EntityID>(r.ids.begin(), r.ids.end());
  }
  std::vector<EntityID> RayIntersectedFirst(TRay const &ray, EntityContainerView entities, double tolerance = BASE_TOLERANCE,
                                            double toleranceIncrement = 0.0,
                                            double maxDistance = std::numeric_limits<double>::max()) const noexcept {
    return RayIntersectedFirst(ray.Origin, ray.Direction, entities, tolerance, toleranceIncrement, maxDistance);
  }
  BatchResult RayIntersectedFirst(std::span<const TRay> rays, double tolerance = BASE_TOLERANCE,
                                  double toleranceIncrement = 0.0,
                                  double maxDistance = std::numeric_limits<double>::max()) const noexcept {
    otg_result *res = nullptr;
    if (!tree_) return empty(rays.size());
    detail::check(otg_ray_first(tree_, rays.size(), reinterpret_cast<const double *>(rays.data()), nullptr, OTG_MEM_HOST,
                                tolerance, toleranceIncrement, maxDistance, &res),
                  "otg_ray_first", ctx().get());
    return take(res);
  }

  // RayIntersectedAll (ot_query.h:1929-2022): every hit, ascending enter distance.
  std::vector<EntityID> RayIntersectedAll(TVector const &rayBasePoint, TVector const &rayHeading, EntityContainerView,
                                          double tolerance = BASE_TOLERANCE, double toleranceIncrement = 0.0,
                                          double maxExaminationDistance = std::numeric_limits<double>::max()) const noexcept {
    TRay ray{rayBasePoint, rayHeading};
    auto r = RayIntersectedAll(std::span<const TRay>(&ray, 1), tolerance, toleranceIncrement, maxExaminationDistance);
    return std::vector<EntityID>(r.ids.begin(), r.ids.end());
  }
  std::vector<EntityID> RayIntersectedAll(TRay const &ray, EntityContainerView entities, double tolerance = BASE_TOLERANCE,
                                          double toleranceIncrement = 0.0,
                                          double maxExaminationDistance = std::numeric_limits<double>::max()) const noexcept {
    return RayIntersectedAll(ray.Origin, ray.Direction, entities, tolerance, toleranceIncrement, maxExaminationDistance);
  }
  BatchResult RayIntersectedAll(std::span<const TRay> rays, double tolerance = BASE_TOLERANCE,
                                double toleranceIncrement = 0.0,
                                double maxExaminationDistance = std::numeric_limits<double>::max()) const noexcept {
    otg_result *res = nullptr;
    if (!tree_) return empty(rays.size());
    detail::check(otg_ray_all(tree_, rays.size(), reinterpret_cast<const double *>(rays.data()), nullptr, OTG_MEM_HOST,
                              tolerance, toleranceIncrement, maxExaminationDistance, &res),
                  "otg_ray_all", ctx().get());
    return take(res);
  }

  // PlaneSearch / PlaneIntersection (ot_query.h:612-688), PlanePositiveSegmentation
  // (:691-753), FrustumCulling (:756-822)
  std::vector<EntityID> PlaneSearch(double distanceOfOrigo, TVector const &planeNormal, EntityContainerView,
                                    double tolerance = BASE_TOLERANCE) const noexcept {
    return planes_one(distanceOfOrigo, planeNormal, OTG_PLANE_SEARCH, tolerance);
  }
  std::vector<EntityID> PlaneSearch(TPlane const &plane, EntityContainerView entities,
                                    double tolerance = BASE_TOLERANCE) const noexcept {
    return PlaneSearch(plane.OrigoDistance, plane.Normal, entities, tolerance);
  }
  std::vector<EntityID> PlaneIntersection(double distanceOfOrigo, TVector const &planeNormal, EntityContainerView entities,
                                          double tolerance = BASE_TOLERANCE) const noexcept {
    return PlaneSearch(distanceOfOrigo, planeNormal, entities, tolerance);
  }
  std::vector<EntityID> PlaneIntersection(TPlane const &plane, EntityContainerView entities,
                                          double tolerance = BASE_TOLERANCE) const noexcept {
    return PlaneSearch(plane, entities, tolerance);
  }
  std::vector<EntityID> PlanePositiveSegmentation(double distanceOfOrigo, TVector const &planeNormal, EntityContainerView,
                                                  double tolerance = BASE_TOLERANCE) const noexcept {
    return planes_one(distanceOfOrigo, planeNormal, OTG_PLANE_POSITIVE, tolerance);
  }
  std::vector<EntityID> PlanePositiveSegmentation(TPlane const &plane, EntityContainerView entities,
                                                  double tolerance = BASE_TOLERANCE) const noexcept {
    return PlanePositiveSegmentation(plane.OrigoDistance, plane.Normal, entities, tolerance);
  }
  std::vector<EntityID> FrustumCulling(std::span<TPlane const> const &boundaryPlanes, EntityContainerView,
                                       double tolerance = BASE_TOLERANCE) const noexcept {
    if (!tree_ || boundaryPlanes.empty()) return {};
    otg_result *res = nullptr;
    detail::check(otg_planes(tree_, 1, (uint32_t)boundaryPlanes.size(), reinterpret_cast<const double *>(boundaryPlanes.data()),
                             nullptr, OTG_MEM_HOST, OTG_PLANE_FRUSTUM, tolerance, &res),
                  "otg_planes", ctx().get());
    BatchResult r = take(res);
    return std::vector<EntityID>(r.ids.begin(), r.ids.end());
  }

  // PickSearch (ot_query.h:426-466)
  std::vector<EntityID> PickSearch(TVector const &pickPoint, EntityContainerView,
                                   double tolerance = BASE_TOLERANCE) const noexcept {
    auto r = PickSearch(std::span<const TVector>(&pickPoint, 1), tolerance);
    return std::vector<EntityID>(r.ids.begin(), r.ids.end());
  }
  BatchResult PickSearch(std::span<const TVector> pickPoints, double tolerance = BASE_TOLERANCE) const noexcept {
    otg_result *res = nullptr;
    if (!tree_) return empty(pickPoints.size());
    detail::check(otg_pick(tree_, pickPoints.size(), reinterpret_cast<const double *>(pickPoints.data()), nullptr,
                           OTG_MEM_HOST, tolerance, &res),
                  "otg_pick", ctx().get());
    return take(res);
  }

  template <typename TExecMode = ParExec>
  std::vector<std::pair<EntityID, EntityID>> CollisionDetection(EntityContainerView, double tolerance = BASE_TOLERANCE,
                                                                TExecMode = {}) const noexcept
    requires(IS_BOX)
  {
    if (!tree_) return {};
    otg_result *res = nullptr;
    detail::check(otg_collide(tree_, 0, tolerance, &res), "otg_collide", ctx().get());
    return pairs(take(res));
  }

  // Two-tree collision detection (ot_query.h:1462-1484): pairs (left id, right id).
  static std::vector<std::pair<EntityID, EntityID>> CollisionDetection(StaticOrthoTreeGPU const &leftTree,
                                                                       EntityContainerView,
                                                                       StaticOrthoTreeGPU const &rightTree,
                                                                       EntityContainerView,
                                                                       double tolerance = BASE_TOLERANCE) noexcept
    requires(IS_BOX)
  {
    if (!leftTree.tree_ || !rightTree.tree_) return {};
    otg_result *res = nullptr;
    detail::check(otg_collide_trees(leftTree.tree_, 0, rightTree.tree_, 0, tolerance, &res), "otg_collide_trees",
                  leftTree.ctx().get());
    return pairs(leftTree.take(res));
  }
  // Member form (ot_query.h:1488-1496): this tree on the left.
  std::vector<std::pair<EntityID, EntityID>> CollisionDetection(EntityContainerView boxes,
                                                                StaticOrthoTreeGPU const &otherTree,
                                                                EntityContainerView otherBoxes,
                                                                double tolerance = BASE_TOLERANCE) const noexcept
    requires(IS_BOX)
  {
    return CollisionDetection(*this, boxes, otherTree, otherBoxes, tolerance);
  }

  // ---- node API (host copy of the device node table, made on first use) ----------------
  std::size_t GetNodeCount() const noexcept { return nodes().depth.size(); }
  depth_t GetMaxDepthID() const noexcept { return nodes().max_depth; }
  static constexpr NodeID GetRootNodeID() noexcept { return 0; }
  std::span<const EntityID> GetNodeEntities(NodeID node) const noexcept {
    const auto &n = nodes();
    return std::span<const EntityID>(n.ids.data() + n.ebegin[node], n.ecount[node]);
  }
  std::pair<NodeID, NodeID> GetNodeChildren(NodeID node) const noexcept {
    const auto &n = nodes();
    return {n.cbegin[node], n.cbegin[node] + n.ccount[node]};
  }
  bool IsNodeEntitiesEmpty(NodeID node) const noexcept { return nodes().ecount[node] == 0; }
  TVector GetNodeMinPoint(NodeID node) const noexcept { return vec(nodes().nmin, node); }
  TVector GetNodeSize(NodeID node) const noexcept { return vec(nodes().nsize, node); }
  TBox GetNodeBox(NodeID node) const noexcept {
    TBox b;
    b.Min = GetNodeMinPoint(node);
    TVector s = GetNodeSize(node);
    for (int d = 0; d < DIM; ++d) b.Max[d] = b.Min[d] + s[d];
    return b;
  }
  depth_t GetNodeDepthID(NodeID node) const noexcept { return nodes().depth[node]; }
  std::span<const EntityID> GetEntityStorage() const noexcept { return nodes().ids; }
  static constexpr bool AreChildNodesOverlapping() noexcept { return IS_BOX; }

  otg_tree *native_handle() const noexcept { return tree_; }

  // Archive of the core in the reference's BinaryOutputArchive layout
  // (serialize(StaticLinearOrthoTreeCore) serialization.h:373-386); largeDataset
  // writes the ORTHOTREE__LARGE_DATASET widths (core/types.h:51-55).
  std::vector<uint8_t> Serialize(bool largeDataset = false) const noexcept {
    std::vector<uint8_t> out;
    if (!tree_) return out;
    const uint32_t flags = largeDataset ? OTG_ARCHIVE_LARGE_DATASET : 0u;
    uint64_t n = 0;
    detail::check(otg_tree_serialize(tree_, 0, flags, nullptr, 0, &n), "otg_tree_serialize", ctx().get());
    out.resize(n);
    detail::check(otg_tree_serialize(tree_, 0, flags, out.data(), n, &n), "otg_tree_serialize", ctx().get());
    return out;
  }

  // serialize(BinaryInputArchive&, StaticLinearOrthoTreeCore&) (serialization.h:373-386):
  // loads an archive written for these entities; the node table and entity order
  // come from the archive.  false for an archive that is malformed or was not
  // written for these entities.
  bool Deserialize(std::span<const uint8_t> archive, EntityContainerView entities, bool largeDataset = false) noexcept {
    return load(archive, reinterpret_cast<const double *>(entities.data()), entities.size(), largeDataset);
  }

  // The archive of an OrthoTreeManaged (serialization.h:433-439) carries the entities
  // after the core: OrthoTreeManaged::Deserialize reads them back.
  bool DeserializeWithEntities(std::span<const uint8_t> archive, bool largeDataset = false) noexcept {
    return load(archive, nullptr, 0, largeDataset);
  }

 private:
  bool load(std::span<const uint8_t> archive, const double *geom, std::size_t n, bool largeDataset) noexcept {
    otg_tree_free(tree_);
    tree_ = nullptr;
    nodes_ = {};
    entity_count_ = 0;
    otg_status st = otg_tree_deserialize(ctx().get(), IS_BOX ? OTG_KIND_BOX : OTG_KIND_POINT, DIM, archive.data(),
                                         archive.size(), largeDataset ? OTG_ARCHIVE_LARGE_DATASET : 0u, n, geom,
                                         OTG_MEM_HOST, &tree_);
    if (st == OTG_ERR_INVALID_ARG) return false;
    detail::check(st, "otg_tree_deserialize", ctx().get());
    otg_set_info si{};
    detail::check(otg_tree_get_set_info(tree_, 0, &si), "otg_tree_get_set_info", ctx().get());
    entity_count_ = si.entity_count;
    return true;
  }

  struct HostNodes {
    bool loaded = false;
    depth_t max_depth = 0;
    std::vector<uint8_t> depth;
    std::vector<uint32_t> ebegin, ecount, cbegin, ccount;
    std::vector<double> nmin, nsize;
    std::vector<EntityID> ids;
  };

  Context &ctx() const noexcept { return ctx_ ? *ctx_ : Context::Default(); }

  const HostNodes &nodes() const noexcept {
    if (!nodes_.loaded && tree_) {
      otg_set_info si{};
      detail::check(otg_tree_get_set_info(tree_, 0, &si), "otg_tree_get_set_info", ctx().get());
      HostNodes &n = nodes_;
      size_t nn = si.node_count;
      n.max_depth = si.max_depth;
      n.depth.resize(nn);
      n.ebegin.resize(nn);
      n.ecount.resize(nn);
      n.cbegin.resize(nn);
      n.ccount.resize(nn);
      n.nmin.resize(nn * DIM);
      n.nsize.resize(nn * DIM);
      n.ids.resize(si.entity_count);
      detail::check(otg_tree_export(tree_, 0, n.depth.data(), n.ebegin.data(), n.ecount.data(), n.cbegin.data(),
                                    n.ccount.data(), nullptr, n.nmin.data(), n.nsize.data(), n.ids.data(), nullptr),
                    "otg_tree_export", ctx().get());
      n.loaded = true;
    }
    return nodes_;
  }

  static TVector vec(const std::vector<double> &a, NodeID node) {
    TVector v;
    for (int d = 0; d < DIM; ++d) v[d] = a[(size_t)node * DIM + d];
    return v;
  }

  std::vector<EntityID> planes_one(double distanceOfOrigo, TVector const &planeNormal, int mode,
                                   double tolerance) const noexcept {
    if (!tree_) return {};
    TPlane pl{distanceOfOrigo, planeNormal};
    otg_result *res = nullptr;
    detail::check(otg_planes(tree_, 1, 1, reinterpret_cast<const double *>(&pl), nullptr, OTG_MEM_HOST, mode, tolerance, &res),
                  "otg_planes", ctx().get());
    BatchResult r = take(res);
    return std::vector<EntityID>(r.ids.begin(), r.ids.end());
  }

  static std::vector<std::pair<EntityID, EntityID>> pairs(BatchResult const &r) {
    std::vector<std::pair<EntityID, EntityID>> out(r.ids.size() / 2);
    for (size_t i = 0; i < out.size(); ++i) out[i] = {r.ids[2 * i], r.ids[2 * i + 1]};
    return out;
  }

  static BatchResult empty(size_t nq) {
    BatchResult r;
    r.offsets.assign(nq + 1, 0);
    return r;
  }

  BatchResult take(otg_result *res) const noexcept {
    BatchResult r;
    uint64_t nq = 0, nids = 0;
    detail::check(otg_result_get_info(res, &nq, &nids), "otg_result_get_info", ctx().get());
    r.offsets.resize(nq + 1);
    r.ids.resize(nids);
    detail::check(otg_result_copy(res, r.offsets.data(), r.ids.data()), "otg_result_copy", ctx().get());
    otg_result_free(res);
    return r;
  }

  Context *ctx_ = nullptr;
  otg_tree *tree_ = nullptr;
  mutable HostNodes nodes_;
  std::size_t entity_count_ = 0;
};

// core/ot_aliases.h:82-91, :225-245
template <int DIM>
using StaticOrthoTreePointND = StaticOrthoTreeGPU<DIM, false>;
namespace detail {
template <int DIM, bool IS_LOOSE_TREE>
struct BoxTreeOf {
  static_assert(IS_LOOSE_TREE, "the GPU path builds loose (LOOSE_FACTOR 2.0) box trees only");
  using type = StaticOrthoTreeGPU<DIM, true>;
};
}  // namespace detail
template <int DIM, bool IS_LOOSE_TREE = true>
using StaticOrthoTreeBoxND = typename detail::BoxTreeOf<DIM, IS_LOOSE_TREE>::type;
using StaticQuadtreePoint = StaticOrthoTreePointND<2>;
using StaticQuadtreeBox = StaticOrthoTreeBoxND<2, true>;
using StaticOctreePoint = StaticOrthoTreePointND<3>;
using StaticOctreeBox = StaticOrthoTreeBoxND<3, true>;
using StaticHexatreePoint = StaticOrthoTreePointND<4>;
using StaticHexatreeBox = StaticOrthoTreeBoxND<4, true>;


// ---- managed (entity-owning) static trees: OrthoTreeManaged (core/ot_managed.h:47-893) ----
// The tree keeps its own copy of the entities (GetData) and the query methods drop the
// entity-span argument, as the reference's managed wrapper does.  Only the static-tree
// members exist (Add / Update / Erase belong to the dynamic core).
template <typename TOrthoTreeCore>
class OrthoTreeManaged {
 public:
  using TVector = typename TOrthoTreeCore::TVector;
  using TBox = typename TOrthoTreeCore::TBox;
  using TRay = typename TOrthoTreeCore::TRay;
  using TPlane = typename TOrthoTreeCore::TPlane;
  using Entity = typename TOrthoTreeCore::Entity;
  using EntityID = typename TOrthoTreeCore::EntityID;
  using EntityContainer = std::vector<Entity>;
  static constexpr std::size_t DEFAULT_TARGET_ELEMENT_NUM_IN_NODES = 16;  // core/configuration.h:83

  OrthoTreeManaged() = default;
  // ot_managed.h:213-316: runtime or compile-time execution tag (SeqExec by default, as
  // ot_managed.h:84-213), with box space
  template <typename TExecMode = SeqExec>
  explicit OrthoTreeManaged(std::span<Entity const> const &entities, std::optional<depth_t> maxDepthID = std::nullopt,
                            std::optional<TBox> boxSpace = std::nullopt,
                            std::size_t maxElementNoInNode = DEFAULT_TARGET_ELEMENT_NUM_IN_NODES,
                            TExecMode execMode = {}) noexcept
      : m_entities(entities.begin(), entities.end()) {
    m_tree.Create(m_entities, maxDepthID, boxSpace, maxElementNoInNode, execMode);
  }
  template <typename TExecMode = SeqExec>
  explicit OrthoTreeManaged(EntityContainer &&entities, std::optional<depth_t> maxDepthID = std::nullopt,
                            std::optional<TBox> boxSpace = std::nullopt,
                            std::size_t maxElementNoInNode = DEFAULT_TARGET_ELEMENT_NUM_IN_NODES,
                            TExecMode execMode = {}) noexcept
      : m_entities(std::move(entities)) {
    m_tree.Create(m_entities, maxDepthID, boxSpace, maxElementNoInNode, execMode);
  }
  template <typename TExecMode>
    requires(std::is_same_v<TExecMode, SeqExec> || std::is_same_v<TExecMode, ParExec>)
  explicit OrthoTreeManaged(TExecMode execMode, std::span<Entity const> const &entities,
                            std::optional<depth_t> maxDepthID = std::nullopt, std::optional<TBox> boxSpace = std::nullopt,
                            std::size_t maxElementNoInNode = DEFAULT_TARGET_ELEMENT_NUM_IN_NODES) noexcept
      : OrthoTreeManaged(entities, maxDepthID, boxSpace, maxElementNoInNode, execMode) {}

  template <typename TExecMode = SeqExec>
  static OrthoTreeManaged Create(std::span<Entity const> const &entities, std::optional<depth_t> maxDepthID = std::nullopt,
                                 std::optional<TBox> boxSpace = std::nullopt,
                                 std::size_t maxElementNoInNode = DEFAULT_TARGET_ELEMENT_NUM_IN_NODES,
                                 TExecMode execMode = {}) noexcept {
    return OrthoTreeManaged(entities, maxDepthID, boxSpace, maxElementNoInNode, execMode);
  }

  TOrthoTreeCore const &GetCore() const noexcept { return m_tree; }

  // serialize(OrthoTreeManaged) serialization.h:433-439: the core, then the
  // entity vector (points as std::array, boxes as BoundingBoxND {Min, Max}).
  std::vector<uint8_t> Serialize(bool largeDataset = false) const noexcept {
    std::vector<uint8_t> out = m_tree.Serialize(largeDataset);
    auto put = [&out](const void *p, std::size_t bytes) {
      const auto *b = static_cast<const uint8_t *>(p);
      out.insert(out.end(), b, b + bytes);
    };
    auto count = [&](uint64_t c) {
      if (largeDataset) {
        put(&c, 8);
      } else {
        // serialized_size_t is uint32 without ORTHOTREE__LARGE_DATASET
        if (c > 0xFFFFFFFFull) detail::fail(OTG_ERR_UNSUPPORTED, "OrthoTreeManaged::Serialize", nullptr);
        const uint32_t c32 = static_cast<uint32_t>(c);
        put(&c32, 4);
      }
    };
    constexpr int D = TOrthoTreeCore::DIMENSION_NO;
    constexpr int ARRAYS = std::is_same_v<Entity, TBox> ? 2 : 1;
    count(m_entities.size());
    for (const Entity &e : m_entities) {
      const double *v = reinterpret_cast<const double *>(&e);
      for (int a = 0; a < ARRAYS; ++a) {
        count(D);
        put(v + a * D, sizeof(double) * D);
      }
    }
    return out;
  }
  // serialize(BinaryInputArchive&, OrthoTreeManaged&) (serialization.h:433-439): the
  // core, then the entity vector it carries.  false for a malformed archive.
  bool Deserialize(std::span<const uint8_t> archive, bool largeDataset = false) noexcept {
    if (!m_tree.DeserializeWithEntities(archive, largeDataset)) return false;
    // the entity vector follows the core: its offset is the core's own length
    std::size_t pos = m_tree.Serialize(largeDataset).size();
    auto take = [&](void *dst, std::size_t bytes) {
      if (archive.size() - pos < bytes) detail::fail(OTG_ERR_INVALID_ARG, "OrthoTreeManaged::Deserialize", nullptr);
      std::memcpy(dst, archive.data() + pos, bytes);
      pos += bytes;
    };
    auto count = [&]() -> uint64_t {
      if (largeDataset) {
        uint64_t c;
        take(&c, 8);
        return c;
      }
      uint32_t c;
      take(&c, 4);
      return c;
    };
    constexpr int D = TOrthoTreeCore::DIMENSION_NO;
    constexpr int ARRAYS = std::is_same_v<Entity, TBox> ? 2 : 1;
    m_entities.assign(count(), Entity{});
    for (Entity &e : m_entities) {
      double *v = reinterpret_cast<double *>(&e);
      for (int a = 0; a < ARRAYS; ++a) {
        (void)count();
        take(v + a * D, sizeof(double) * D);
      }
    }
    return true;
  }
  EntityContainer const &GetData() const noexcept { return m_entities; }
  Entity const &Get(EntityID entityID) const noexcept { return m_entities[entityID]; }

  std::vector<EntityID> PickSearch(TVector const &pickPoint, double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.PickSearch(pickPoint, m_entities, tolerance);
  }
  std::vector<EntityID> RangeSearch(TBox const &range, RangeSearchMode mode = RangeSearchMode::Inside,
                                    double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.RangeSearch(range, m_entities, mode, tolerance);
  }
  std::vector<EntityID> PlaneSearch(double distanceOfOrigo, TVector const &planeNormal,
                                    double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.PlaneSearch(distanceOfOrigo, planeNormal, m_entities, tolerance);
  }
  std::vector<EntityID> PlaneSearch(TPlane const &plane, double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.PlaneSearch(plane, m_entities, tolerance);
  }
  std::vector<EntityID> PlaneIntersection(double distanceOfOrigo, TVector const &planeNormal,
                                          double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.PlaneIntersection(distanceOfOrigo, planeNormal, m_entities, tolerance);
  }
  std::vector<EntityID> PlaneIntersection(TPlane const &plane, double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.PlaneIntersection(plane, m_entities, tolerance);
  }
  std::vector<EntityID> PlanePositiveSegmentation(double distanceOfOrigo, TVector const &planeNormal,
                                                  double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.PlanePositiveSegmentation(distanceOfOrigo, planeNormal, m_entities, tolerance);
  }
  std::vector<EntityID> PlanePositiveSegmentation(TPlane const &plane, double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.PlanePositiveSegmentation(plane, m_entities, tolerance);
  }
  std::vector<EntityID> FrustumCulling(std::span<TPlane const> const &boundaryPlanes,
                                       double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.FrustumCulling(boundaryPlanes, m_entities, tolerance);
  }
  std::vector<EntityID> GetNearestNeighbors(TVector const &pt, std::size_t k,
                                            double maxDistanceWithin = std::numeric_limits<double>::max(),
                                            double tolerance = BASE_TOLERANCE) const noexcept {
    return m_tree.GetNearestNeighbors(pt, k, maxDistanceWithin, m_entities, tolerance);
  }
  template <typename TExecMode = ParExec>
  std::vector<std::pair<EntityID, EntityID>> CollisionDetection(double tolerance = BASE_TOLERANCE,
                                                                TExecMode execMode = {}) const noexcept {
    return m_tree.CollisionDetection(m_entities, tolerance, execMode);
  }
  std::vector<std::pair<EntityID, EntityID>> CollisionDetection(OrthoTreeManaged const &otherTree,
                                                                double tolerance = BASE_TOLERANCE) const noexcept {
    return TOrthoTreeCore::CollisionDetection(m_tree, m_entities, otherTree.m_tree, otherTree.m_entities, tolerance);
  }
  static std::vector<std::pair<EntityID, EntityID>> CollisionDetection(OrthoTreeManaged const &leftTree,
                                                                       OrthoTreeManaged const &rightTree,
                                                                       double tolerance = BASE_TOLERANCE) noexcept {
    return leftTree.CollisionDetection(rightTree, tolerance);
  }
  std::vector<EntityID> RayIntersectedAll(TVector const &rayBasePoint, TVector const &rayHeading,
                                          double tolerance = BASE_TOLERANCE, double toleranceIncrement = 0.0,
                                          double maxDistance = std::numeric_limits<double>::max()) const noexcept {
    return m_tree.RayIntersectedAll(rayBasePoint, rayHeading, m_entities, tolerance, toleranceIncrement, maxDistance);
  }
  std::vector<EntityID> RayIntersectedAll(TRay const &ray, double tolerance = BASE_TOLERANCE,
                                          double toleranceIncrement = 0.0,
                                          double maxDistance = std::numeric_limits<double>::max()) const noexcept {
    return RayIntersectedAll(ray.Origin, ray.Direction, tolerance, toleranceIncrement, maxDistance);
  }
  std::vector<EntityID> RayIntersectedFirst(TVector const &rayBasePoint, TVector const &rayHeading,
                                            double tolerance = BASE_TOLERANCE, double toleranceIncrement = 0.0,
                                            double maxDistance = std::numeric_limits<double>::max()) const noexcept {
    return m_tree.RayIntersectedFirst(rayBasePoint, rayHeading, m_entities, tolerance, toleranceIncrement, maxDistance);
  }
  std::vector<EntityID> RayIntersectedFirst(TRay const &ray, double tolerance = BASE_TOLERANCE,
                                            double toleranceIncrement = 0.0,
                                            double maxDistance = std::numeric_limits<double>::max()) const noexcept {
    return RayIntersectedFirst(ray.Origin, ray.Direction, tolerance, toleranceIncrement, maxDistance);
  }

 private:
  TOrthoTreeCore m_tree;
  EntityContainer m_entities;
};

// core/ot_aliases.h:417-447
template <int DIM>
using StaticTreePointManagedND = OrthoTreeManaged<StaticOrthoTreePointND<DIM>>;
template <int DIM, bool IS_LOOSE_TREE = true>
using StaticTreeBoxManagedND = OrthoTreeManaged<StaticOrthoTreeBoxND<DIM, IS_LOOSE_TREE>>;
using StaticQuadtreePointM = StaticTreePointManagedND<2>;
using StaticQuadtreeBoxM = StaticTreeBoxManagedND<2, true>;
using StaticOctreePointM = StaticTreePointManagedND<3>;
using StaticOctreeBoxM = StaticTreeBoxManagedND<3, true>;

}  // namespace otg
