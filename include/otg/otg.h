/*
 * otg.h -- C ABI of libotg, the MI355X (gfx950) static linear ortho-tree.
 *
 * Plain C types only: pointers, sizes, status codes.  No exceptions cross this
 * boundary.  Every entry point returns an otg_status (0 = OK).
 *
 * The reference (attcs/Octree, header-only C++20) has no FFI; its seam is the
 * Core template parameter of OrthoTreeQueryBase (core/ot_query.h:98-99) and the
 * static-tree aliases (core/ot_aliases.h:82-91, :237, :240).  Each entry point
 * below names the reference member it replaces.  The C++ drop-in
 * (include/otg/orthotree_gpu.hpp) maps these onto the reference's class and
 * method names.
 *
 * Geometry layouts (all FP64, row-major "AoS", identical to a
 * std::span<const std::array<double, D>> / BoundingBoxND<D> span):
 *   point   : D doubles
 *   box     : D min doubles, then D max doubles
 *   ray     : D origin doubles, then D direction doubles
 * Entity ids are the positions in the caller's span (ContiguousIndex strategy,
 * core/entity_adapter.h:170-174); in a forest they are local to each set.
 */
#ifndef OTG_OTG_H
#define OTG_OTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define OTG_API __attribute__((visibility("default")))
#else
#define OTG_API
#endif

typedef int otg_status;
#define OTG_OK 0
#define OTG_ERR_INVALID_ARG 1   /* bad dimension, cap == 0, depth out of range, NULL handle */
#define OTG_ERR_HIP 2           /* a HIP runtime call failed */
#define OTG_ERR_OOM 3           /* device allocation failed (reference: std::bad_alloc) */
#define OTG_ERR_UNSUPPORTED 4   /* input the GPU path does not mirror (see otg_status_string) */
#define OTG_ERR_OUT_OF_SPACE 5  /* a point outside an explicit box: the reference Build never ends */
#define OTG_ERR_KEY_OVERFLOW 6  /* set-id + location bits of a forest exceed 64 */
#define OTG_ERR_NO_DEVICE 7     /* no HIP device visible */
#define OTG_ERR_CAPACITY 8      /* a caller's output buffer is too small (the size needed is reported) */

#define OTG_MEM_HOST 0   /* pointer is host memory */
#define OTG_MEM_DEVICE 1 /* pointer is device memory of the context's device */

#define OTG_KIND_POINT 0 /* StaticOrthoTreePointND: loose 1.0, MinPoint geometry */
#define OTG_KIND_BOX 1   /* StaticOrthoTreeBoxND<D, true>: loose 2.0, MBR geometry */

#define OTG_RANGE_INSIDE 0  /* RangeSearchMode::Inside  (core/types.h) */
#define OTG_RANGE_OVERLAP 1 /* RangeSearchMode::Overlap */

#define OTG_PLANE_SEARCH 0   /* PlaneSearch / PlaneIntersection: entities the plane hits */
#define OTG_PLANE_POSITIVE 1 /* PlanePositiveSegmentation: entities not on the negative side */
#define OTG_PLANE_FRUSTUM 2  /* FrustumCulling: entities on no boundary plane's negative side */

typedef struct otg_ctx otg_ctx;
typedef struct otg_tree otg_tree;
typedef struct otg_result otg_result;
typedef struct otg_knn_plan otg_knn_plan;

typedef struct {
  int32_t kind;        /* OTG_KIND_* */
  int32_t dim;         /* D */
  uint32_t nsets;      /* 1 for a tree built with otg_tree_build */
  uint32_t reserved;
  uint64_t entity_count;
  uint64_t node_count; /* over all sets */
  uint64_t device_bytes;
} otg_tree_info;

typedef struct {
  uint64_t entity_offset; /* first entity of the set in the forest input */
  uint64_t entity_count;
  uint64_t node_offset;   /* first node of the set in the forest node table */
  uint64_t node_count;    /* GetNodeCount() of the set's tree */
  uint32_t max_depth;     /* GetMaxDepthID() */
  uint32_t max_elements;  /* GetMaxElementNum() */
  double space_min[16];   /* tree box (GridSpaceIndexing box space) */
  double space_max[16];
} otg_set_info;

/* ---- context ------------------------------------------------------------ */
/* One HIP device + one default-priority stream + a caching device allocator. */
OTG_API otg_status otg_ctx_create(int device, otg_ctx **out);
OTG_API void otg_ctx_destroy(otg_ctx *ctx);
OTG_API otg_status otg_ctx_synchronize(otg_ctx *ctx);
OTG_API const char *otg_status_string(otg_status status);
/* Last error detail text of this context (empty if none). */
OTG_API const char *otg_ctx_last_error(const otg_ctx *ctx);
/* Device memory held by the context's allocator cache, and a cache release. */
OTG_API otg_status otg_ctx_mem_stats(const otg_ctx *ctx, uint64_t *bytes_in_use, uint64_t *bytes_cached);
OTG_API otg_status otg_ctx_release_cache(otg_ctx *ctx);
/* Per-kernel timing with HIP events on the context stream (for benchmarks). */
OTG_API otg_status otg_ctx_set_profiling(otg_ctx *ctx, int enable);
/* Copies up to cap entries; returns the number of distinct kernels seen. */
OTG_API otg_status otg_ctx_kernel_times(otg_ctx *ctx, uint32_t cap, char *names /* cap x 64 */, double *total_ms,
                                        uint64_t *launches, uint32_t *count);
OTG_API otg_status otg_ctx_reset_kernel_times(otg_ctx *ctx);

/* ---- build -------------------------------------------------------------- */
/* StaticLinearOrthoTreeCore::Create<ParExec> (core/ot_static_linear_core.h:523-599)
 * and the PAR_EXEC constructor (:118-129).  box_or_null: 2*dim doubles (host) or
 * NULL for the entities' bounding box (internal_geometry_module.h:342-370);
 * max_depth 0 = EstimateMaxDepth (detail/utils.h:613-622). */
OTG_API otg_status otg_tree_build(otg_ctx *ctx, int kind, int dim, uint64_t n, const double *geom, int mem,
                                  const double *box_or_null, uint32_t max_depth, uint32_t max_elements, otg_tree **out);

/* Create<SeqExec> (ot_static_linear_core.h:523-599 with ARE_LOCATIONS_SORTED = false):
 * the same tree as otg_tree_build, numbered as the reference's unsorted build
 * numbers it (Build<false> :417-519): each node's entities are split by repeated
 * two-pointer partitions (detail/partitioning.h:215-241) in the order they arrive,
 * so children follow their pivots' order and entities the partitions' order.  The
 * reference's default execution mode.  otg_tree_export's sorted_keys are then the
 * location keys in that storage order. */
OTG_API otg_status otg_tree_build_seq(otg_ctx *ctx, int kind, int dim, uint64_t n, const double *geom, int mem,
                                      const double *box_or_null, uint32_t max_depth, uint32_t max_elements,
                                      otg_tree **out);

/* Batched build: nsets independent trees of uneven sizes built together.
 * set_offsets: nsets+1 host values (entity ranges in geom); boxes_or_null:
 * nsets x 2*dim host doubles; max_depths_or_null: nsets (0 = estimate);
 * max_elements: nsets.  Each set is the tree otg_tree_build would make. */
OTG_API otg_status otg_forest_build(otg_ctx *ctx, int kind, int dim, uint32_t nsets, const uint64_t *set_offsets,
                                    const double *geom, int mem, const double *boxes_or_null,
                                    const uint32_t *max_depths_or_null, const uint32_t *max_elements, otg_tree **out);

OTG_API void otg_tree_free(otg_tree *tree);
OTG_API otg_status otg_tree_get_info(const otg_tree *tree, otg_tree_info *out);
OTG_API otg_status otg_tree_get_set_info(const otg_tree *tree, uint32_t set, otg_set_info *out);

/* Node table of one set (for parity and the node API, :178-268).  Any pointer
 * may be NULL.  Node ids are BFS ranks local to the set; children of node i are
 * [child_begin[i], child_begin[i] + child_count[i]) (GetNodeChildren :191-202);
 * entity segment [entity_begin, +entity_count) indexes entity_ids
 * (GetNodeEntities :204-212, m_entityStorage).  node_min / node_size: node_count
 * x dim (GetNodeMinPoint / GetNodeSize).  node_key: Morton node key with
 * sentinel bit (si_morton.h:216-226).  sorted_keys: location codes of the
 * entity storage order (code << dbits | depth, dbits = 0 for points). */
OTG_API otg_status otg_tree_export(const otg_tree *tree, uint32_t set, uint8_t *node_depth, uint32_t *entity_begin,
                                   uint32_t *entity_count, uint32_t *child_begin, uint32_t *child_count,
                                   uint64_t *node_key, double *node_min, double *node_size, uint32_t *entity_ids,
                                   uint64_t *sorted_keys);

/* Static-core archive of one set: the byte stream the reference's
 * BinaryOutputArchive writes for a StaticLinearOrthoTreeCore (serialize of the
 * core serialization.h:373-386 and of its base :359-371; std::variant
 * serialization/stl/variant.h, std::vector stl/vector.h, std::array
 * stl/array.h; BinaryArchive::operator& serialization/binary_archive.h:85-175,
 * little-endian).  The node storage alternative is the one Create picks
 * (ot_static_linear_core.h:330-344, :367-372, :573-578).  flags:
 * OTG_ARCHIVE_LARGE_DATASET writes sizes and NodeStorageGeneral segments as
 * uint64 (ORTHOTREE__LARGE_DATASET, core/types.h:51-55 and :172-176).
 * *size receives the archive length; the bytes are written only when buf is
 * not NULL, and buf_size must then be at least that length.  The archive of an
 * OrthoTreeManaged (serialization.h:433-439) is this followed by the entity
 * vector, which the owner of the entities appends. */
#define OTG_ARCHIVE_LARGE_DATASET 1u
OTG_API otg_status otg_tree_serialize(const otg_tree *tree, uint32_t set, uint32_t flags, uint8_t *buf,
                                      uint64_t buf_size, uint64_t *size);

/* Loads a static-core archive into a tree on ctx: serialize(BinaryInputArchive&,
 * StaticLinearOrthoTreeCore&) (serialization.h:359-386) of the bytes
 * otg_tree_serialize writes (flags as there).  geom = the tree's n entities in id
 * order (memory space mem), or NULL for the archive of an OrthoTreeManaged, whose
 * entity vector follows the core (serialization.h:433-439).  The node table, node
 * geometry and entity order are taken from the archive, not rebuilt; it is checked
 * to be a static linear tree (children after their parent and deeper, each node's
 * entities followed by its children's, ids a permutation) whose entities lie in
 * their nodes, else OTG_ERR_INVALID_ARG.  Queries then answer as on the archived
 * tree; point trees with overfull max-depth leaves get the finer kNN / range index
 * a build makes.  A loaded point tree has no location keys (the archive holds
 * none): otg_tree_export of sorted_keys returns OTG_ERR_UNSUPPORTED for it. */
OTG_API otg_status otg_tree_deserialize(otg_ctx *ctx, int kind, int dim, const uint8_t *buf, uint64_t size,
                                        uint32_t flags, uint64_t n, const double *geom, int mem, otg_tree **out);

/* ---- autotuning ---------------------------------------------------------- */
/* Settings of the batched path.  leaf_capacity is the build's maxElementNumInNode
 * (ot_static_linear_core.h:523-529), which the caller passes to otg_forest_build;
 * the others live in the context and apply to every later build and query on it:
 * index_capacity = leaf capacity of the finer kNN / range index (0: the tree's),
 * waves_per_simd = occupancy cap of the kNN walks (0: as registers allow),
 * order_queries = 1 to walk batches of >= 16,384 queries in Morton order,
 * persistent = 1 to run the kNN walks as a persistent launch, range_records = the
 * records a RangeSearch wavefront count walk keeps per query for its fill (0: a
 * batch's record budget over its capped queries, 256 to 8192).  Query results never
 * depend on them. */
typedef struct {
  uint32_t leaf_capacity;
  uint32_t index_capacity;
  uint32_t waves_per_simd;
  uint32_t order_queries;
  uint32_t persistent; /* 1: kNN walks as a persistent launch pulling 64-query tiles across sets */
  uint32_t range_records; /* kept records per capped RangeSearch query (0: from the batch) */
  double step_ms; /* measured build + queries with these settings (otg_autotune) */
} otg_tuning;

/* No counterpart in the reference (its leaf capacity is a caller's constant):
 * measures the forest build of the given sets (as otg_forest_build, leaf capacity
 * from {8, 16, 32, 64}) followed by the kNN batch (k, default tolerance) and/or the
 * RangeSearch batch (Inside), over the candidate settings, and returns the fastest.
 * The pick is cached per workload shape for the process and applied to ctx.  Either
 * query batch may be empty (nq 0); qmem is the queries' memory space. */
OTG_API otg_status otg_autotune(otg_ctx *ctx, int kind, int dim, uint32_t nsets, const uint64_t *set_offsets,
                                const double *geom, int mem, const double *boxes_or_null,
                                const uint32_t *max_depths_or_null, uint64_t nq_knn, const double *knn_points,
                                const uint32_t *knn_sets, uint32_t k, uint64_t nq_range, const double *ranges,
                                const uint32_t *range_sets, int qmem, otg_tuning *out);
/* Applies index_capacity, waves_per_simd and order_queries to the context (the
 * untuned defaults are 0, 0, 1), or reads them back. */
OTG_API otg_status otg_ctx_set_tuning(otg_ctx *ctx, const otg_tuning *tuning);
OTG_API otg_status otg_ctx_get_tuning(const otg_ctx *ctx, otg_tuning *tuning);

/* ---- queries ------------------------------------------------------------ */
/* Results are CSR: offsets[nq+1] (uint64) and ids (uint32, local entity ids).
 * set_ids: per-query set of a forest (NULL = set 0); same memory space as the
 * queries.  Every set id must be below the tree's set count (OTG_ERR_INVALID_ARG
 * otherwise, checked before any query runs).  A batch whose result would hold
 * 2^32 - 1 or more ids is rejected with OTG_ERR_UNSUPPORTED (split the batch). */

/* OrthoTreeQueryBase::GetNearestNeighbors(point, k, maxDistanceWithin, entities,
 * tolerance) (core/ot_query.h:1208-1270).  max_within = DBL_MAX for the
 * overload without a distance limit.  Ids sorted by ascending distance.  Any k:
 * above 64 the final radius comes from a k-entry heap per query in global memory. */
OTG_API otg_status otg_knn(const otg_tree *tree, uint64_t nq, const double *points, const uint32_t *set_ids, int mem,
                           uint32_t k, double max_within, double tolerance, otg_result **out);

/* Measurement helper, no counterpart in the reference: runs the kNN batch and sums the
 * minimal set it must touch -- out[0] = visits of nodes whose wall distance is below the
 * query's final radius, out[1] = entities held by those nodes, out[2] = result ids. */
OTG_API otg_status otg_knn_footprint(const otg_tree *tree, uint64_t nq, const double *points, const uint32_t *set_ids,
                                     int mem, uint32_t k, double max_within, double tolerance, uint64_t *out);

/* RangeSearch(range, entities, mode, tolerance) (core/ot_query.h:476-583). */
OTG_API otg_status otg_range(const otg_tree *tree, uint64_t nq, const double *ranges, const uint32_t *set_ids, int mem,
                             int mode, double tolerance, otg_result **out);

/* RayIntersectedFirst(origin, direction, entities, tolerance, toleranceIncrement,
 * maxDistance) (core/ot_query.h:2035-2146).  Ids in descending enter distance. */
OTG_API otg_status otg_ray_first(const otg_tree *tree, uint64_t nq, const double *rays, const uint32_t *set_ids, int mem,
                                 double tolerance, double tolerance_increment, double max_distance, otg_result **out);

/* CollisionDetection<ParExec>(entities, tolerance) of one set of a box tree
 * (core/ot_query.h:1901-1906).  Result: nq = pair count, ids = 2*pairs
 * (i < j), sorted by (i, j). */
OTG_API otg_status otg_collide(const otg_tree *tree, uint32_t set, double tolerance, otg_result **out);

/* Static CollisionDetection(leftTree, leftEntities, rightTree, rightEntities,
 * tolerance) (core/ot_query.h:1462-1484) and its member form (:1488-1496), for
 * one set of each box tree (same dimension, same device).  Result: nq = pair
 * count, ids = 2*pairs (left id, right id), sorted by (left, right). */
OTG_API otg_status otg_collide_trees(const otg_tree *left, uint32_t left_set, const otg_tree *right,
                                     uint32_t right_set, double tolerance, otg_result **out);

/* PickSearch(pickPoint, entities, tolerance) (core/ot_query.h:426-466): boxes that
 * contain the point, or points within tolerance of it.  tolerance >= 0. */
OTG_API otg_status otg_pick(const otg_tree *tree, uint64_t nq, const double *points, const uint32_t *set_ids, int mem,
                            double tolerance, otg_result **out);

/* PlaneSearch(distanceOfOrigo, planeNormal, entities, tolerance) (core/ot_query.h:612-636),
 * PlanePositiveSegmentation (:691-741) and FrustumCulling(boundaryPlanes, entities,
 * tolerance) (:756-822).  planes: nq x nplanes x (1 + dim) doubles, each plane its origo
 * distance then its unit normal (PlaneND, adapters/general.h:364-368); nplanes must be 1 for
 * OTG_PLANE_SEARCH and OTG_PLANE_POSITIVE; a frustum with no planes gives nothing. */
OTG_API otg_status otg_planes(const otg_tree *tree, uint64_t nq, uint32_t nplanes, const double *planes,
                              const uint32_t *set_ids, int mem, int mode, double tolerance, otg_result **out);

/* RayIntersectedAll(origin, direction, entities, tolerance, toleranceIncrement,
 * maxExaminationDistance) (core/ot_query.h:1929-2022): every entity the ray hits
 * entering within max_distance (0 = no limit, as the reference), ascending enter
 * distance (ties by id). */
OTG_API otg_status otg_ray_all(const otg_tree *tree, uint64_t nq, const double *rays, const uint32_t *set_ids, int mem,
                               double tolerance, double tolerance_increment, double max_distance, otg_result **out);

/* Allocation-free kNN.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same ids in the same order as otg_knn, for
 * device-resident points (and set_ids, or NULL), written into the caller's device
 * CSR: offsets[nq + 1] and at most ids_capacity ids; *nids receives the id count,
 * also with OTG_ERR_CAPACITY (then it is the size needed and ids is untouched).
 * Every temporary lives in a fixed 96 MiB scratch array per device, held until the
 * process ends, so a call takes at most 262,144 queries (OTG_ERR_INVALID_ARG
 * beyond) with k <= 64; a batch whose tolerance bands need more than the array
 * holds fails with OTG_ERR_UNSUPPORTED.  Calls on one device share its array and
 * are serialised by a process-wide lock (the array keeps the last call's contents).
 * Queries are walked in input order; the call synchronises the tree's stream. */
OTG_API otg_status otg_knn_noalloc(const otg_tree *tree, uint64_t nq, const double *points, const uint32_t *set_ids,
                                   uint32_t k, double max_within, double tolerance, uint64_t *offsets, uint32_t *ids,
                                   uint64_t ids_capacity, uint64_t *nids);

/* kNN plans.  No counterpart in the reference (its queries are single calls): for
 * callers that send many batches of the same size, e.g. a solver that queries the
 * same particles every step.  otg_knn_plan_create records one otg_knn batch over
 * device-resident points[nq * dim] (and set_ids, or NULL) as a HIP graph, with
 * every buffer allocated up front; otg_knn_plan_run replays it on the current
 * contents of those buffers (the caller may rewrite them between runs, not move
 * them) and synchronises.  Results are the ids otg_knn returns for the same
 * contents, in the same order.  The tuning in force at creation is kept.
 * otg_knn_plan_result gives the last run's device CSR, valid until the next run or
 * otg_knn_plan_free.  1 <= k <= 64 (OTG_ERR_UNSUPPORTED beyond); a set id outside
 * the tree fails the run with OTG_ERR_INVALID_ARG.  Free plans before their tree. */
OTG_API otg_status otg_knn_plan_create(const otg_tree *tree, uint64_t nq, const double *points,
                                       const uint32_t *set_ids, uint32_t k, double max_within, double tolerance,
                                       otg_knn_plan **out);
OTG_API otg_status otg_knn_plan_run(otg_knn_plan *plan);
OTG_API otg_status otg_knn_plan_result(const otg_knn_plan *plan, const uint64_t **offsets, const uint32_t **ids,
                                       uint64_t *nids);
/* Host copy of the last run's CSR: offsets[nq + 1], ids[nids]. */
OTG_API otg_status otg_knn_plan_copy(const otg_knn_plan *plan, uint64_t *offsets, uint32_t *ids);
OTG_API void otg_knn_plan_free(otg_knn_plan *plan);

/* ---- build plans: a forest build replayed as a HIP graph ----
 * No counterpart in the reference (a header-only library that builds on the host):
 * for callers that rebuild the same sets every step from new coordinates (the
 * batched path's "build + query" loop).  otg_build_plan_create builds the point
 * forest otg_forest_build would make from device-resident points (every set's box
 * given), then records that build as a HIP graph with every buffer allocated up
 * front and node capacities of the current contents plus an eighth and 4,096;
 * otg_build_plan_run replays it on the current contents of the same points buffer
 * (rewritten in place, not moved) and synchronises.  The plan's tree
 * (otg_build_plan_tree; freed with the plan, never by otg_tree_free) then equals a
 * fresh otg_forest_build of those contents and takes every query.  New contents that
 * need more nodes than the plan holds fail the run with OTG_ERR_CAPACITY (the tree is
 * then left without nodes; make a new plan); a point outside its set's box fails it
 * with OTG_ERR_OUT_OF_SPACE.  Free kNN and step plans made on the tree first. */
typedef struct otg_build_plan otg_build_plan;
OTG_API otg_status otg_build_plan_create(otg_ctx *ctx, int dim, uint32_t nsets, const uint64_t *set_offsets,
                                         const double *points, const double *boxes, const uint32_t *max_depths_or_null,
                                         const uint32_t *max_elements, otg_build_plan **out);
OTG_API otg_status otg_build_plan_run(otg_build_plan *plan);
OTG_API otg_status otg_build_plan_tree(const otg_build_plan *plan, const otg_tree **tree);
OTG_API void otg_build_plan_free(otg_build_plan *plan);

/* Range plans: one RangeSearch batch (otg_range) of device-resident boxes on a 2-D or
 * 3-D tree, recorded as a HIP graph with every buffer allocated up front; the id
 * capacity comes from the batch's result at creation (ids + ids / 8 + 65,536).
 * otg_range_plan_run replays it on the current contents of the boxes (and set ids)
 * and synchronises; results equal otg_range's for the same contents.  A result that
 * outgrows the plan fails the run with OTG_ERR_CAPACITY (no ids written; make a new
 * plan); a set id outside the tree with OTG_ERR_INVALID_ARG.  No counterpart in the
 * reference, like the kNN plans.  Free plans before their tree. */
typedef struct otg_range_plan otg_range_plan;
OTG_API otg_status otg_range_plan_create(const otg_tree *tree, uint64_t nq, const double *ranges,
                                         const uint32_t *set_ids, int mode, double tolerance, otg_range_plan **out);
OTG_API otg_status otg_range_plan_run(otg_range_plan *plan);
OTG_API otg_status otg_range_plan_result(const otg_range_plan *plan, const uint64_t **offsets, const uint32_t **ids,
                                         uint64_t *nids);
/* Host copy of the last run's CSR: offsets[nq + 1], ids[nids]. */
OTG_API otg_status otg_range_plan_copy(const otg_range_plan *plan, uint64_t *offsets, uint32_t *ids);
OTG_API void otg_range_plan_free(otg_range_plan *plan);

/* A build plan and query plans made on its tree, replayed together: the build, then
 * the kNN batch (knn_or_null) and the range batch (range_or_null); with both, the
 * range batch runs on a second stream of the step plan beside the kNN batch.
 * otg_step_plan_run synchronises; afterwards the tree is current and each query
 * plan's result / copy give its batch's ids, as after otg_build_plan_run followed by
 * each plan's own run.  Free it before the plans. */
typedef struct otg_step_plan otg_step_plan;
OTG_API otg_status otg_step_plan_create(otg_build_plan *build, otg_knn_plan *knn_or_null,
                                        otg_range_plan *range_or_null, otg_step_plan **out);
OTG_API otg_status otg_step_plan_run(otg_step_plan *step);
OTG_API void otg_step_plan_free(otg_step_plan *step);

OTG_API otg_status otg_result_get_info(const otg_result *res, uint64_t *nq, uint64_t *nids);
OTG_API otg_status otg_result_copy(const otg_result *res, uint64_t *offsets, uint32_t *ids);
/* Device-resident CSR (valid until otg_result_free). */
OTG_API otg_status otg_result_device(const otg_result *res, const uint64_t **offsets, const uint32_t **ids);
OTG_API void otg_result_free(otg_result *res);

/* ---- serving: resident trees and continuous batching ----
 * No counterpart in the reference (a header-only library): the runtime for callers that keep
 * many trees resident in HBM and send streams of query batches.  Trees are built once and
 * stay resident under `hbm_budget_bytes`; adding one that does not fit evicts the least
 * recently used idle trees.  `nworkers` worker threads, each with its own stream, take all
 * pending requests, merge requests on the same tree with equal parameters into one launch and
 * split the results back.  A request's result equals the direct call's (otg_knn / otg_range /
 * otg_ray_first) on the same queries, whatever it was batched with. */
typedef struct otg_server otg_server;
OTG_API otg_status otg_server_create(int device, uint64_t hbm_budget_bytes, uint32_t nworkers, otg_server **out);
/* No counterpart in the reference: the most queries a worker takes into one batch
 * (default 2,097,152; a single larger request still forms a batch alone).  Smaller
 * batches shorten the service time a request waits for and lower the tail latency at
 * a given load, at the price of more launches and less throughput.  Applies to batches
 * formed after the call. */
OTG_API otg_status otg_server_set_batch_limit(otg_server *server, uint64_t max_queries);
/* Waits for running batches; frees every resident tree and unclaimed result. */
OTG_API void otg_server_destroy(otg_server *server);
/* Registers the tree under tree_id (keeping a host copy of the entities, so that an evicted
 * tree is rebuilt when a request needs it) and builds it as otg_tree_build does; replaces a
 * tree registered under the same id once nothing uses it. */
OTG_API otg_status otg_server_add_tree(otg_server *server, uint32_t tree_id, int kind, int dim, uint64_t n,
                                       const double *geom, int mem, const double *box_or_null, uint32_t max_depth,
                                       uint32_t max_elements);
OTG_API otg_status otg_server_remove_tree(otg_server *server, uint32_t tree_id);
/* *present: 0 not registered, 1 resident, 2 registered but evicted. */
OTG_API otg_status otg_server_has_tree(otg_server *server, uint32_t tree_id, int *present);
/* query 0 = kNN (k, p0 = max_within, p1 = tolerance); 1 = range (mode, p0 = tolerance);
 * 2 = ray first (p0 = tolerance, p1 = tolerance increment, p2 = max distance).  The queries
 * (host memory) are copied; returns at once with a request id. */
OTG_API otg_status otg_server_submit(otg_server *server, uint32_t tree_id, int query, uint64_t nq,
                                     const double *queries, uint32_t k, double p0, double p1, double p2, int mode,
                                     uint64_t *request_id);
/* Blocks until the request is served; *out is a host-resident result owned by the caller
 * (otg_result_copy / otg_result_free; otg_result_device is not available for it).
 * latency_ms = submit to done, queue_ms = submit to batch start (either may be NULL). */
OTG_API otg_status otg_server_wait(otg_server *server, uint64_t request_id, otg_result **out, double *latency_ms,
                                   double *queue_ms);
/* otg_server_wait with where the time went: detail[0] latency, [1] queueing (submit to
 * batch start, including the wait that gathers requests for an evicted tree), [2] the
 * batch's service time, [3] of which making trees resident (rebuilds, waits for
 * memory), [4] requests in the batch, [5] 1 if the request's tree was not resident
 * when the batch formed (all in ms). */
OTG_API otg_status otg_server_wait_detail(otg_server *server, uint64_t request_id, otg_result **out, double *detail);
/* stats[15] = batches, launches, requests served, queries served, evictions, resident bytes,
 * resident trees, pending requests, rebuilds of evicted trees, registered trees, and
 * microseconds spent building trees, in query launches overall, and within those staging
 * the inputs, in the launch itself, and splitting the results per request. */
OTG_API otg_status otg_server_stats(otg_server *server, uint64_t *stats);

#ifdef __cplusplus
}
#endif
#endif /* OTG_OTG_H */
