/*
 * otg_oracle.c -- CPU restatement of OrthoTree's static linear-tree build and
 * query path (attcs/Octree v3.2.4), written in plain C11.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * It is never linked into libotg and never used as a fallback.
 *
 * Every function follows the reference file:line cited in its comment, with the
 * same floating-point operation order (compile with -ffp-contract=off, no
 * fast-math).  Where the reference leaves an order unspecified (std::sort tie
 * order, heap tie order) the oracle picks a deterministic one (entity id).
 *
 * Pinning (how each part is checked independently of this restatement):
 *   encode            Morton KATs general.tests.cpp:136-217 + closed form bit i*D+d
 *   node keys         general.tests.cpp:465-494 KATs + closed form
 *   quantisation      closed form special cases (c<=0, c==size, clamp) in tests
 *   node table        structural invariant checker written independently in
 *                     Python (tests/node_table_rules.py) on seeded inputs
 *   kNN               brute force with the closed-form set rule + KATs
 *                     general.tests.cpp:2475-2803
 *   range             naive scan (random.tests.cpp:220-246 restated)
 *   collision         brute force (general.tests.cpp:4695-4705) + Issue38 procedure
 *   ray first         KATs general.tests.cpp:1280-1543 + closed-form candidate rule
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OC_API __attribute__((visibility("default")))
#define OC_MAXD 16

/* adapters/general.h:57  BASE_TOLERANCE = epsilon * 100 */
#define OC_BASE_TOL (DBL_EPSILON * 100.0)

/* std::max / std::min semantics (return first argument unless the second is
 * strictly greater / smaller), as used throughout the reference. */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

static unsigned bit_width64(uint64_t x) { return x ? 64u - (unsigned)__builtin_clzll(x) : 0u; }

/* ------------------------------------------------------------------------- */
/* Tree record                                                                */
/* ------------------------------------------------------------------------- */
typedef struct {
  int dim;
  int is_box;
  uint32_t max_depth;
  uint32_t cap;
  uint32_t n;
  double space_min[OC_MAXD];
  double space_max[OC_MAXD];
  double size[OC_MAXD];      /* GridSpaceIndexing::m_sizeInDimensions */
  double raster[OC_MAXD];    /* rasterizerFactors */
  double deraster[OC_MAXD];  /* derasterizerFactors */
  double *depth_size;        /* (max_depth+2) x dim: OrthoTreeCoreBase::m_nodeSize */

  uint32_t node_count;
  uint32_t node_alloc;
  uint8_t *node_depth;
  uint32_t *node_ebegin, *node_ecount;  /* own entity segment */
  uint32_t *node_rbegin, *node_rend;    /* whole subtree entity range */
  uint32_t *node_cbegin, *node_ccount;  /* children */
  uint64_t *node_key;                   /* Morton node key with sentinel bit */
  double *node_min;                     /* node_count x dim */
  double *node_size;                    /* node_count x dim (MBR only) */

  uint32_t *entity_ids;  /* m_entityStorage */
  uint64_t *sorted_code; /* location codes in entity-storage order */
  uint32_t *sorted_depth;
  const double *geom;    /* entity geometry, borrowed (AoS: D or 2D doubles) */
  int status;            /* 0 ok, 1 out of space (boxes), 2 unsupported input */
} oc_tree;

/* ------------------------------------------------------------------------- */
/* Morton space indexing                                                      */
/* ------------------------------------------------------------------------- */

/* si_morton.h:498-574 Encode: bit i of gridID[d] -> bit i*D + d.  The pdep,
 * Part1By2 and generic-loop variants are all this mapping. */
static uint64_t oc_encode(const uint32_t *g, int D) {
  uint64_t code = 0;
  for (int i = 0; i < 32; ++i)
    for (int d = 0; d < D; ++d)
      if ((g[d] >> i) & 1u) {
        unsigned sh = (unsigned)(i * D + d);
        if (sh < 64) code |= (uint64_t)1 << sh;
      }
  return code;
}

/* si_morton.h:576-622 Decode(nodeKey, maxDepth): grid id of the node's min
 * cell, i.e. the key's digits placed at levels maxDepth-depth .. maxDepth-1. */
static void oc_decode(uint64_t node_key, uint32_t depth, uint32_t max_depth, int D, uint32_t *g) {
  for (int d = 0; d < D; ++d) g[d] = 0;
  unsigned shift = 0;
  for (uint32_t lvl = max_depth - depth; lvl < max_depth; ++lvl)
    for (int d = 0; d < D; ++d, ++shift) g[d] |= (uint32_t)((node_key >> shift) & 1u) << lvl;
}

/* si_morton.h:216-226 GetNodeID(location, maxDepth) */
static uint64_t oc_node_key(uint64_t code, uint32_t depth, uint32_t max_depth, int D) {
  return ((uint64_t)1 << (depth * (unsigned)D)) | (code >> ((max_depth - depth) * (unsigned)D));
}

OC_API uint64_t oc_morton_encode(const uint32_t *g, int D) { return oc_encode(g, D); }
OC_API uint64_t oc_morton_node_key(uint64_t code, uint32_t depth, uint32_t max_depth, int D) {
  return oc_node_key(code, depth, max_depth, D);
}
OC_API void oc_morton_decode(uint64_t key, uint32_t depth, uint32_t max_depth, int D, uint32_t *g) {
  oc_decode(key, depth, max_depth, D, g);
}

/* utils.h:613-622 EstimateMaxDepth<D, MAX_THEORETICAL_DEPTH_ID> */
static uint32_t oc_max_theoretical_depth(int D) {
  /* si_morton.h:75-77: (bits(NodeID) - 1) / D; 32-bit NodeID for D <= 2. */
  return D <= 2 ? (uint32_t)(31 / D) : (uint32_t)(63 / D);
}
OC_API uint32_t oc_estimate_max_depth(uint64_t n, uint64_t cap, int D) {
  if (n <= cap) return 2;
  uint64_t nleaf = n / cap;
  uint32_t est = (uint32_t)(log2((double)nleaf) / (double)D);
  uint32_t hi = oc_max_theoretical_depth(D);
  if (est < 2) est = 2;
  if (est > hi) est = hi;
  return est;
}

/* si_grid.h:65-103 GridSpaceIndexing ctor; ot_base.h:121-134 node sizes */
static void oc_init_space(oc_tree *t, const double *bmin, const double *bmax) {
  int D = t->dim;
  double res = (double)((uint64_t)1 << t->max_depth);
  for (int d = 0; d < D; ++d) {
    t->space_min[d] = bmin[d];
    t->space_max[d] = bmax[d];
    t->size[d] = bmax[d] - bmin[d];
    t->deraster[d] = t->size[d] / res;
    t->raster[d] = (t->size[d] == 0) ? 1.0 : (res / t->size[d]);
  }
  uint32_t nd = t->max_depth + 2;
  t->depth_size = (double *)malloc(sizeof(double) * nd * (size_t)D);
  double loose = t->is_box ? 2.0 : 1.0; /* configuration.h:111-115 LOOSE_FACTOR */
  for (uint32_t k = 0; k < nd; ++k, loose *= 0.5)
    for (int d = 0; d < D; ++d) t->depth_size[k * D + d] = t->size[d] * loose;
}

/* si_grid.h:145-175 GetPointGridID + si_mortongrid.h:57-76 GetLocation(point).
 * Out-of-space points get Location(0, 0) (ALLOW_OUT_OF_SPACE_INSERTION). */
static void oc_point_location(const oc_tree *t, const double *p, uint64_t *code, uint32_t *depth) {
  int D = t->dim;
  uint32_t g[OC_MAXD];
  uint32_t max_raster_id = (uint32_t)(((uint64_t)1 << t->max_depth) - 1);
  for (int d = 0; d < D; ++d) {
    double c = p[d] - t->space_min[d];
    if (c < -OC_BASE_TOL || c > t->size[d] + OC_BASE_TOL) {
      *code = 0;
      *depth = 0;
      return;
    }
    if (c <= 0) g[d] = 0;
    else {
      uint32_t r = (uint32_t)(c * t->raster[d]);
      g[d] = r < max_raster_id ? r : max_raster_id;
    }
  }
  *code = oc_encode(g, D);
  *depth = t->max_depth;
}

/* si_grid.h:221-285 GetLooseBoxGridData<2.0> + si_mortongrid.h:78-97.
 * internal_geometry_module.h:214-226 DoesRangeContainBoxAD (IGM range, TBox). */
static void oc_box_location(oc_tree *t, const double *b, uint64_t *code, uint32_t *depth) {
  int D = t->dim;
  for (int d = 0; d < D; ++d) {
    double rmin = t->space_min[d], rmax = t->space_max[d];
    double bmin = b[d], bmax = b[D + d];
    if (rmin - OC_BASE_TOL > bmin || bmin > rmax + OC_BASE_TOL ||
        rmin - OC_BASE_TOL > bmax || bmax > rmax + OC_BASE_TOL) {
      *code = 0;
      *depth = 0;
      return;
    }
  }
  uint32_t g[OC_MAXD];
  double max_rel = 0.0;
  for (int d = 0; d < D; ++d) {
    double center = (b[d] + b[D + d]) * 0.5;
    g[d] = (uint32_t)((center - t->space_min[d]) * t->raster[d]);
  }
  for (int d = 0; d < D; ++d) {
    double sz = b[D + d] - b[d];
    max_rel = smax(max_rel, sz * t->raster[d]);
  }
  uint32_t mrg = (uint32_t)ceil(max_rel);
  uint32_t level = 0;
  if (mrg != 0) level = bit_width64((uint64_t)(mrg - 1));
  if (level > t->max_depth) t->status = 2;
  for (int d = 0; d < D; ++d)
    if (g[d] >> t->max_depth) t->status = 2;
  uint64_t c = oc_encode(g, D);
  unsigned sh = level * (unsigned)D;
  if (sh < 64) c = (c >> sh) << sh;
  else c = 0;
  *code = c;
  *depth = t->max_depth - level;
}

/* ------------------------------------------------------------------------- */
/* Build                                                                      */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint64_t code;
  uint32_t depth;
  uint32_t id;
} oc_loc;

/* si_morton.h:85-143 Location::operator< (code, then depth); ties by id. */
static int oc_loc_cmp(const void *a, const void *b) {
  const oc_loc *x = (const oc_loc *)a, *y = (const oc_loc *)b;
  if (x->code != y->code) return x->code < y->code ? -1 : 1;
  if (x->depth != y->depth) return x->depth < y->depth ? -1 : 1;
  if (x->id != y->id) return x->id < y->id ? -1 : 1;
  return 0;
}

static void oc_reserve_nodes(oc_tree *t, uint32_t need) {
  if (need <= t->node_alloc) return;
  uint32_t na = t->node_alloc ? t->node_alloc * 2 : 64;
  while (na < need) na *= 2;
  int D = t->dim;
  t->node_depth = (uint8_t *)realloc(t->node_depth, na);
  t->node_ebegin = (uint32_t *)realloc(t->node_ebegin, na * 4u);
  t->node_ecount = (uint32_t *)realloc(t->node_ecount, na * 4u);
  t->node_rbegin = (uint32_t *)realloc(t->node_rbegin, na * 4u);
  t->node_rend = (uint32_t *)realloc(t->node_rend, na * 4u);
  t->node_cbegin = (uint32_t *)realloc(t->node_cbegin, na * 4u);
  t->node_ccount = (uint32_t *)realloc(t->node_ccount, na * 4u);
  t->node_key = (uint64_t *)realloc(t->node_key, na * 8u);
  t->node_min = (double *)realloc(t->node_min, (size_t)na * D * 8u);
  if (t->is_box) t->node_size = (double *)realloc(t->node_size, (size_t)na * D * 8u);
  t->node_alloc = na;
}

/* ot_static_linear_core.h:346-368 InitNodeGeometry (MinPoint):
 * center = ((g + 2^(M-depth)*0.5) * deraster + spaceMin)  (si_grid.h:124-143,
 * si_mortongrid.h:123-126), minPoint = center - nodeSize[depth+1]. */
static void oc_node_minpoint(const oc_tree *t, uint64_t key, uint32_t depth, double *out) {
  int D = t->dim;
  uint32_t g[OC_MAXD];
  oc_decode(key, depth, t->max_depth, D, g);
  double half_grid = (double)((uint64_t)1 << (t->max_depth - depth)) * 0.5;
  for (int d = 0; d < D; ++d) {
    double center = ((double)g[d] + half_grid) * t->deraster[d] + t->space_min[d];
    out[d] = center - t->depth_size[(depth + 1) * D + d];
  }
}

/* ot_static_linear_core.h:383-412 CreateNode (children recorded as a range) */
static uint32_t oc_create_node(oc_tree *t, uint64_t code, uint32_t depth, uint32_t ebegin, uint32_t ecount,
                               uint32_t rbegin, uint32_t rend) {
  uint32_t id = t->node_count;
  oc_reserve_nodes(t, id + 1);
  t->node_count = id + 1;
  t->node_depth[id] = (uint8_t)depth;
  t->node_ebegin[id] = ebegin;
  t->node_ecount[id] = ecount;
  t->node_rbegin[id] = rbegin;
  t->node_rend[id] = rend;
  t->node_cbegin[id] = 0;
  t->node_ccount[id] = 0;
  t->node_key[id] = oc_node_key(code, depth, t->max_depth, t->dim);
  if (!t->is_box) oc_node_minpoint(t, t->node_key[id], depth, &t->node_min[(size_t)id * t->dim]);
  return id;
}

typedef struct {
  uint64_t code;
  uint32_t depth;
  uint32_t begin, length;
} oc_qitem;

/* ot_static_linear_core.h:417-519 Build<ARE_LOCATIONS_SORTED = true>:
 * BFS queue; leaf iff count <= cap or depth >= maxDepth; boxes keep the
 * non-refinable prefix (entity depth == node depth); children grouped by the
 * digit at depth+1 (ChildCheckerFixedDepth si_morton.h:176-191) with the
 * path-compressed LCA depth (LowestCommonAncestorCalculator :346-387). */
static void oc_build_bfs(oc_tree *t, const oc_loc *locs, uint32_t n) {
  int D = t->dim;
  uint32_t M = t->max_depth;
  uint64_t child_mask = ((uint64_t)1 << D) - 1;
  oc_qitem *queue = (oc_qitem *)malloc(sizeof(oc_qitem) * (size_t)(2 * n + 16));
  size_t qh = 0, qt = 0;
  queue[qt++] = (oc_qitem){0, 0, 0, n};
  /* children of the node currently being split are the next queue entries;
   * record (parent id, first child queue slot) to fill child ranges once the
   * children are created (node id == queue position: BFS). */
  uint32_t *qparent = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(2 * n + 16));
  qparent[0] = UINT32_MAX;
  while (qh < qt) {
    oc_qitem it = queue[qh];
    uint32_t parent = qparent[qh];
    ++qh;
    uint32_t depth = it.depth;
    uint32_t node_id;
    if (it.length <= t->cap || depth >= M) {
      node_id = oc_create_node(t, it.code, depth, it.begin, it.length, it.begin, it.begin + it.length);
    } else {
      uint32_t b = it.begin, e = it.begin + it.length;
      uint32_t nr = 0;
      int leaf_all_nr = 0;
      if (t->is_box) {
        while (b + nr < e && locs[b + nr].depth == depth) ++nr; /* partition_point */
        if (nr == it.length) leaf_all_nr = 1;
      }
      if (leaf_all_nr) {
        node_id = oc_create_node(t, it.code, depth, it.begin, it.length, it.begin, e);
      } else {
        uint32_t examined = M - (depth + 1); /* GetExaminationLevelID(depth+1) */
        uint64_t mask = child_mask << (examined * (unsigned)D);
        uint32_t cb = b + nr;
        uint32_t first_child_slot = (uint32_t)qt;
        uint32_t nchild = 0;
        while (cb < e) {
          const oc_loc *pivot = &locs[cb];
          uint64_t flag = pivot->code & mask;
          uint32_t min_depth = pivot->depth;
          uint64_t diff = 0;
          uint32_t ce = cb;
          while (ce < e && (locs[ce].code & mask) == flag) {
            if (t->is_box && locs[ce].depth < min_depth) min_depth = locs[ce].depth;
            diff |= pivot->code ^ locs[ce].code;
            ++ce;
          }
          uint32_t level_diff = (bit_width64(diff) + (unsigned)D - 1) / (unsigned)D;
          uint32_t lca = M - level_diff;
          if (min_depth < lca) lca = min_depth;
          queue[qt] = (oc_qitem){pivot->code, lca, cb, ce - cb};
          qparent[qt] = UINT32_MAX - 1; /* filled below */
          ++qt;
          ++nchild;
          cb = ce;
        }
        node_id = oc_create_node(t, it.code, depth, it.begin, nr, it.begin, e);
        for (uint32_t k = 0; k < nchild; ++k) qparent[first_child_slot + k] = node_id;
        /* BFS: the children receive node ids equal to their queue slots. */
        t->node_cbegin[node_id] = first_child_slot;
        t->node_ccount[node_id] = nchild;
      }
    }
    (void)parent;
  }
  free(queue);
  free(qparent);
}

/* ot_static_linear_core.h:280-327 InitializeMinimalNodeGeometry (MBR), bottom-up
 * by depth: union of own entity boxes and children's (min, min+size); stored as
 * (min, max-min).  Children always have a larger depth than their parent. */
static void oc_build_mbr(oc_tree *t) {
  int D = t->dim;
  for (int dep = (int)t->max_depth; dep >= 0; --dep) {
    for (uint32_t id = 0; id < t->node_count; ++id) {
      if (t->node_depth[id] != (uint32_t)dep) continue;
      double mn[OC_MAXD], mx[OC_MAXD];
      for (int d = 0; d < D; ++d) {
        mn[d] = DBL_MAX;
        mx[d] = -DBL_MAX;
      }
      for (uint32_t j = t->node_ebegin[id]; j < t->node_ebegin[id] + t->node_ecount[id]; ++j) {
        const double *bx = &t->geom[(size_t)t->entity_ids[j] * 2 * D];
        for (int d = 0; d < D; ++d) {
          mn[d] = smin(mn[d], bx[d]);
          mx[d] = smax(mx[d], bx[D + d]);
        }
      }
      for (uint32_t c = t->node_cbegin[id]; c < t->node_cbegin[id] + t->node_ccount[id]; ++c) {
        for (int d = 0; d < D; ++d) {
          double cmin = t->node_min[(size_t)c * D + d];
          double cmax = cmin + t->node_size[(size_t)c * D + d]; /* GetNodeBox :253-268 */
          mn[d] = smin(mn[d], cmin);
          mx[d] = smax(mx[d], cmax);
        }
      }
      for (int d = 0; d < D; ++d) {
        t->node_min[(size_t)id * D + d] = mn[d];
        t->node_size[(size_t)id * D + d] = mx[d] - mn[d];
      }
    }
  }
}

/* internal_geometry_module.h:342-370 GetBoundingBoxAD */
static void oc_bounding_box(const double *geom, uint32_t n, int D, int is_box, double *bmin, double *bmax) {
  for (int d = 0; d < D; ++d) {
    bmin[d] = DBL_MAX;
    bmax[d] = -DBL_MAX;
  }
  size_t stride = is_box ? 2 * (size_t)D : (size_t)D;
  for (uint32_t i = 0; i < n; ++i) {
    const double *g = &geom[i * stride];
    for (int d = 0; d < D; ++d) {
      bmin[d] = smin(bmin[d], g[d]);
      bmax[d] = smax(bmax[d], is_box ? g[D + d] : g[d]);
    }
  }
}

OC_API void oc_bounding_box_export(const double *geom, uint32_t n, int D, int is_box, double *bmin, double *bmax) {
  oc_bounding_box(geom, n, D, is_box, bmin, bmax);
}

/* Parallel encode + sort of the PAR path (ot_static_linear_core.h:549-572: the
 * transform and the sort run par_unseq; Build<true> stays serial): each thread
 * encodes and sorts one chunk, then sorted chunks are merged pairwise, a round of
 * merges at a time.  The comparator is a total order (ids break ties), so the
 * result equals the serial sort's. */
typedef struct {
  const oc_tree *t;
  const double *geom;
  oc_loc *locs;
  uint32_t lo, hi;
  int status;
} oc_enc_job;

static void *oc_enc_run(void *arg) {
  oc_enc_job *j = (oc_enc_job *)arg;
  oc_tree local = *j->t; /* private header: the location functions record status in it */
  local.status = 0;
  const int D = local.dim;
  for (uint32_t i = j->lo; i < j->hi; ++i) {
    j->locs[i].id = i;
    if (local.is_box) oc_box_location(&local, &j->geom[(size_t)i * 2 * D], &j->locs[i].code, &j->locs[i].depth);
    else oc_point_location(&local, &j->geom[(size_t)i * D], &j->locs[i].code, &j->locs[i].depth);
    if (!local.is_box && j->locs[i].depth == 0) local.status = 2;
  }
  qsort(j->locs + j->lo, j->hi - j->lo, sizeof(oc_loc), oc_loc_cmp);
  j->status = local.status;
  return NULL;
}

typedef struct {
  const oc_loc *src;
  oc_loc *dst;
  uint32_t lo, mid, hi;
} oc_merge_job;

static void *oc_merge_run(void *arg) {
  oc_merge_job *m = (oc_merge_job *)arg;
  uint32_t a = m->lo, b = m->mid, o = m->lo;
  while (a < m->mid && b < m->hi) m->dst[o++] = (oc_loc_cmp(&m->src[b], &m->src[a]) < 0) ? m->src[b++] : m->src[a++];
  while (a < m->mid) m->dst[o++] = m->src[a++];
  while (b < m->hi) m->dst[o++] = m->src[b++];
  return NULL;
}

/* locs[0..n) encoded and sorted with `threads` threads; returns the location status. */
static int oc_encode_sort_mt(const oc_tree *t, const double *geom, oc_loc *locs, uint32_t n, int threads) {
  if (threads < 1) threads = 1;
  if ((uint32_t)threads > n) threads = n ? (int)n : 1;
  oc_enc_job *jobs = (oc_enc_job *)calloc((size_t)threads, sizeof(oc_enc_job));
  pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  uint32_t *bounds = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)threads + 1));
  for (int i = 0; i <= threads; ++i) bounds[i] = (uint32_t)(((uint64_t)n * (uint64_t)i) / (uint64_t)threads);
  for (int i = 0; i < threads; ++i) {
    jobs[i] = (oc_enc_job){t, geom, locs, bounds[i], bounds[i + 1], 0};
    if (threads > 1) pthread_create(&th[i], NULL, oc_enc_run, &jobs[i]);
    else oc_enc_run(&jobs[i]);
  }
  int status = 0;
  for (int i = 0; i < threads; ++i) {
    if (threads > 1) pthread_join(th[i], NULL);
    if (jobs[i].status) status = jobs[i].status;
  }
  /* pairwise merges of sorted runs */
  oc_loc *tmp = threads > 1 ? (oc_loc *)malloc(sizeof(oc_loc) * (size_t)n) : NULL;
  oc_loc *src = locs, *dst = tmp;
  int runs = threads;
  uint32_t *rb = bounds;
  while (runs > 1) {
    const int pairs = runs / 2;
    oc_merge_job *mj = (oc_merge_job *)calloc((size_t)pairs + 1, sizeof(oc_merge_job));
    for (int p2 = 0; p2 < pairs; ++p2) {
      mj[p2] = (oc_merge_job){src, dst, rb[2 * p2], rb[2 * p2 + 1], rb[2 * p2 + 2]};
      pthread_create(&th[p2], NULL, oc_merge_run, &mj[p2]);
    }
    if (runs & 1) { /* the odd run is copied over */
      mj[pairs] = (oc_merge_job){src, dst, rb[runs - 1], rb[runs], rb[runs]};
      oc_merge_run(&mj[pairs]);
    }
    for (int p2 = 0; p2 < pairs; ++p2) pthread_join(th[p2], NULL);
    free(mj);
    int nr = 0;
    for (int r = 0; r < runs; r += 2) rb[nr++] = rb[r];
    rb[nr] = n;
    runs = nr;
    oc_loc *x = src;
    src = dst;
    dst = x;
  }
  if (src != locs) memcpy(locs, src, sizeof(oc_loc) * (size_t)n);
  free(tmp);
  free(bounds);
  free(th);
  free(jobs);
  return status;
}

OC_API oc_tree *oc_build_mt(int D, int is_box, uint32_t n, const double *geom, const double *box_or_null,
                            uint32_t max_depth_in, uint32_t cap, int threads);

/* ot_static_linear_core.h:523-599 Create<ParExec>: box (given or bounding),
 * depth (given or EstimateMaxDepth), encode every entity, sort by location,
 * BFS build, MBR pass.  Returns NULL on invalid parameters. */
OC_API oc_tree *oc_build(int D, int is_box, uint32_t n, const double *geom, const double *box_or_null,
                         uint32_t max_depth_in, uint32_t cap) {
  return oc_build_mt(D, is_box, n, geom, box_or_null, max_depth_in, cap, 1);
}

OC_API oc_tree *oc_build_mt(int D, int is_box, uint32_t n, const double *geom, const double *box_or_null,
                            uint32_t max_depth_in, uint32_t cap, int threads) {
  if (D < 1 || D > OC_MAXD || cap == 0) return NULL;
  oc_tree *t = (oc_tree *)calloc(1, sizeof(oc_tree));
  t->dim = D;
  t->is_box = is_box;
  t->cap = cap;
  t->n = n;
  t->geom = geom;
  double bmin[OC_MAXD], bmax[OC_MAXD];
  if (box_or_null) {
    for (int d = 0; d < D; ++d) {
      bmin[d] = box_or_null[d];
      bmax[d] = box_or_null[D + d];
    }
  } else {
    oc_bounding_box(geom, n, D, is_box, bmin, bmax);
  }
  t->max_depth = max_depth_in ? max_depth_in : oc_estimate_max_depth(n, cap, D);
  if (t->max_depth < 1 || t->max_depth > oc_max_theoretical_depth(D) || t->max_depth > 19) {
    free(t);
    return NULL;
  }
  oc_init_space(t, bmin, bmax);
  if (n == 0) return t;

  oc_loc *locs = (oc_loc *)malloc(sizeof(oc_loc) * n);
  /* an out-of-space point sets status 2: the reference Build never ends for it */
  const int st = oc_encode_sort_mt(t, geom, locs, n, threads);
  if (st) t->status = st;
  if (t->status == 2) {
    free(locs);
    return t;
  }
  t->entity_ids = (uint32_t *)malloc(4u * n);
  t->sorted_code = (uint64_t *)malloc(8u * n);
  t->sorted_depth = (uint32_t *)malloc(4u * n);
  for (uint32_t i = 0; i < n; ++i) {
    t->entity_ids[i] = locs[i].id;
    t->sorted_code[i] = locs[i].code;
    t->sorted_depth[i] = locs[i].depth;
  }
  oc_build_bfs(t, locs, n);
  free(locs);
  if (is_box) oc_build_mbr(t);
  return t;
}

/* ------------------------------------------------------------------------- */
/* SEQ path: Create<SeqExec> with ARE_LOCATIONS_SORTED = false                */
/* ------------------------------------------------------------------------- */
/* detail/partitioning.h:215-241 Partitioning::partition, the two-pointer scheme
 * (std::partition of the non-refinable boxes, ot_static_linear_core.h:455-456, is
 * the same scheme in libstdc++'s bidirectional __partition).  Predicate: entity
 * depth == depth (by_depth), else (code & mask) == flag.  Returns the end of the
 * front part. */
static uint32_t oc_partition(oc_loc *a, uint32_t first, uint32_t last, int by_depth, uint32_t depth, uint64_t mask,
                             uint64_t flag) {
#define OC_P(x) (by_depth ? ((x).depth == depth) : (((x).code & mask) == flag))
  for (;;) {
    while (first != last && OC_P(a[first])) ++first;
    if (first == last) return first;
    for (;;) {
      if (first == --last) return first;
      if (OC_P(a[last])) break;
    }
    oc_loc tmp = a[first];
    a[first] = a[last];
    a[last] = tmp;
    ++first;
  }
#undef OC_P
}

/* The partition on (code, id) arrays, for checking the scheme against the C++
 * library's std::partition. */
OC_API uint32_t oc_partition_codes(uint64_t *codes, uint32_t *ids, uint32_t n, uint64_t mask, uint64_t flag) {
  oc_loc *a = (oc_loc *)malloc(sizeof(oc_loc) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) a[i] = (oc_loc){codes[i], 0, ids[i]};
  uint32_t m = oc_partition(a, 0, n, 0, 0, mask, flag);
  for (uint32_t i = 0; i < n; ++i) {
    codes[i] = a[i].code;
    ids[i] = a[i].id;
  }
  free(a);
  return m;
}

/* ot_static_linear_core.h:417-519 Build<false>: the BFS queue of Build<true>, but
 * a node's range is split in the order its entities have reached: the boxes at
 * the node's depth are partitioned to the front (:449-459), then repeatedly the
 * first entity of the rest is the pivot and its child's entities are partitioned
 * to the front (:484-500), each child queued with the LCA of its entities. */
static void oc_build_bfs_seq(oc_tree *t, oc_loc *locs, uint32_t n) {
  int D = t->dim;
  uint32_t M = t->max_depth;
  uint64_t child_mask = ((uint64_t)1 << D) - 1;
  oc_qitem *queue = (oc_qitem *)malloc(sizeof(oc_qitem) * (size_t)(2 * n + 16));
  size_t qh = 0, qt = 0;
  queue[qt++] = (oc_qitem){0, 0, 0, n};
  while (qh < qt) {
    oc_qitem it = queue[qh++];
    uint32_t depth = it.depth;
    if (it.length <= t->cap || depth >= M) {
      oc_create_node(t, it.code, depth, it.begin, it.length, it.begin, it.begin + it.length);
      continue;
    }
    uint32_t b = it.begin, e = it.begin + it.length;
    uint32_t nr = 0;
    if (t->is_box) {
      nr = oc_partition(locs, b, e, 1, depth, 0, 0) - b;
      if (nr == it.length) {
        oc_create_node(t, it.code, depth, it.begin, it.length, it.begin, e);
        continue;
      }
    }
    uint32_t examined = M - (depth + 1);
    uint64_t mask = child_mask << (examined * (unsigned)D);
    uint32_t cb = b + nr;
    uint32_t first_child_slot = (uint32_t)qt, nchild = 0;
    while (cb < e) {
      const oc_loc pivot = locs[cb];
      uint32_t ce = oc_partition(locs, cb, e, 0, 0, mask, pivot.code & mask);
      uint32_t min_depth = pivot.depth;
      uint64_t diff = 0;
      for (uint32_t i = cb; i < ce; ++i) { /* LowestCommonAncestorCalculator (si_morton.h:346-387) */
        if (t->is_box && locs[i].depth < min_depth) min_depth = locs[i].depth;
        diff |= pivot.code ^ locs[i].code;
      }
      uint32_t lca = M - (bit_width64(diff) + (unsigned)D - 1) / (unsigned)D;
      if (min_depth < lca) lca = min_depth;
      queue[qt++] = (oc_qitem){pivot.code, lca, cb, ce - cb};
      ++nchild;
      cb = ce;
    }
    uint32_t id = oc_create_node(t, it.code, depth, it.begin, nr, it.begin, e);
    t->node_cbegin[id] = first_child_slot; /* BFS: node id == queue slot */
    t->node_ccount[id] = nchild;
  }
  free(queue);
}

/* ot_static_linear_core.h:523-599 Create<SeqExec>: locations in entity order, no
 * sort, Build<false>, MBR pass. */
OC_API oc_tree *oc_build_seq(int D, int is_box, uint32_t n, const double *geom, const double *box_or_null,
                             uint32_t max_depth_in, uint32_t cap) {
  if (D < 1 || D > OC_MAXD || cap == 0) return NULL;
  oc_tree *t = (oc_tree *)calloc(1, sizeof(oc_tree));
  t->dim = D;
  t->is_box = is_box;
  t->cap = cap;
  t->n = n;
  t->geom = geom;
  double bmin[OC_MAXD], bmax[OC_MAXD];
  if (box_or_null) {
    for (int d = 0; d < D; ++d) {
      bmin[d] = box_or_null[d];
      bmax[d] = box_or_null[D + d];
    }
  } else {
    oc_bounding_box(geom, n, D, is_box, bmin, bmax);
  }
  t->max_depth = max_depth_in ? max_depth_in : oc_estimate_max_depth(n, cap, D);
  if (t->max_depth < 1 || t->max_depth > oc_max_theoretical_depth(D) || t->max_depth > 19) {
    free(t);
    return NULL;
  }
  oc_init_space(t, bmin, bmax);
  if (n == 0) return t;
  oc_loc *locs = (oc_loc *)malloc(sizeof(oc_loc) * n);
  for (uint32_t i = 0; i < n; ++i) {
    locs[i].id = i;
    if (is_box) oc_box_location(t, &geom[(size_t)i * 2 * D], &locs[i].code, &locs[i].depth);
    else oc_point_location(t, &geom[(size_t)i * D], &locs[i].code, &locs[i].depth);
    if (!is_box && locs[i].depth == 0) t->status = 2;
  }
  if (t->status == 2) {
    free(locs);
    return t;
  }
  oc_build_bfs_seq(t, locs, n);
  t->entity_ids = (uint32_t *)malloc(4u * n);
  t->sorted_code = (uint64_t *)malloc(8u * n);
  t->sorted_depth = (uint32_t *)malloc(4u * n);
  for (uint32_t i = 0; i < n; ++i) {
    t->entity_ids[i] = locs[i].id;
    t->sorted_code[i] = locs[i].code;
    t->sorted_depth[i] = locs[i].depth;
  }
  free(locs);
  if (is_box) oc_build_mbr(t);
  return t;
}

OC_API void oc_tree_free(oc_tree *t) {
  if (!t) return;
  free(t->depth_size);
  free(t->node_depth);
  free(t->node_ebegin);
  free(t->node_ecount);
  free(t->node_rbegin);
  free(t->node_rend);
  free(t->node_cbegin);
  free(t->node_ccount);
  free(t->node_key);
  free(t->node_min);
  free(t->node_size);
  free(t->entity_ids);
  free(t->sorted_code);
  free(t->sorted_depth);
  free(t);
}

OC_API int oc_tree_status(const oc_tree *t) { return t->status; }
OC_API uint32_t oc_tree_node_count(const oc_tree *t) { return t->node_count; }
OC_API uint32_t oc_tree_max_depth(const oc_tree *t) { return t->max_depth; }

/* Export the node table; any pointer may be NULL. */
OC_API void oc_tree_export(const oc_tree *t, uint8_t *depth, uint32_t *ebegin, uint32_t *ecount, uint32_t *cbegin,
                           uint32_t *ccount, uint64_t *key, double *nmin, double *nsize, uint32_t *entity_ids,
                           uint64_t *sorted_code, uint32_t *sorted_depth, double *space) {
  uint32_t nc = t->node_count;
  int D = t->dim;
  if (depth) memcpy(depth, t->node_depth, nc);
  if (ebegin) memcpy(ebegin, t->node_ebegin, 4u * nc);
  if (ecount) memcpy(ecount, t->node_ecount, 4u * nc);
  if (cbegin) memcpy(cbegin, t->node_cbegin, 4u * nc);
  if (ccount) memcpy(ccount, t->node_ccount, 4u * nc);
  if (key) memcpy(key, t->node_key, 8u * nc);
  if (nmin) memcpy(nmin, t->node_min, (size_t)8 * nc * D);
  if (nsize) {
    for (uint32_t i = 0; i < nc; ++i)
      for (int d = 0; d < D; ++d)
        nsize[(size_t)i * D + d] = t->is_box ? t->node_size[(size_t)i * D + d] : t->depth_size[t->node_depth[i] * D + d];
  }
  if (entity_ids && t->entity_ids) memcpy(entity_ids, t->entity_ids, 4u * t->n);
  if (sorted_code && t->sorted_code) memcpy(sorted_code, t->sorted_code, 8u * t->n);
  if (sorted_depth && t->sorted_depth) memcpy(sorted_depth, t->sorted_depth, 4u * t->n);
  if (space) {
    for (int d = 0; d < D; ++d) {
      space[d] = t->space_min[d];
      space[D + d] = t->space_max[d];
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Geometry predicates                                                        */
/* ------------------------------------------------------------------------- */
static inline const double *oc_nmin(const oc_tree *t, uint32_t id) { return &t->node_min[(size_t)id * t->dim]; }
static inline const double *oc_nsize(const oc_tree *t, uint32_t id) {
  return t->is_box ? &t->node_size[(size_t)id * t->dim] : &t->depth_size[t->node_depth[id] * t->dim];
}

/* internal_geometry_module.h:424-458 GetBoxWallDistanceAD */
static double oc_wall_distance(const double *q, const double *mn, const double *sz, int D, int inside_zero) {
  double cd[OC_MAXD], hs[OC_MAXD];
  int inside = 1;
  for (int d = 0; d < D; ++d) {
    hs[d] = sz[d] * 0.5;
    cd[d] = fabs(mn[d] + hs[d] - q[d]);
    inside &= cd[d] <= hs[d];
  }
  if (inside) {
    if (inside_zero) return 0.0;
    double m = hs[0];
    for (int d = 0; d < D; ++d) m = smin(m, hs[d] - cd[d]);
    return m;
  }
  double s2 = 0.0;
  for (int d = 0; d < D; ++d) {
    double v = smax(0.0, cd[d] - hs[d]);
    s2 += v * v;
  }
  return sqrt(s2);
}

/* adapters/general.h:135-146 Distance(ptL, ptR) */
static inline double oc_distance(const double *a, const double *b, int D) {
  double d2 = 0.0;
  for (int d = 0; d < D; ++d) {
    double v = a[d] - b[d];
    d2 += v * v;
  }
  return sqrt(d2);
}

/* internal_geometry_module.h:148-161 AreBoxesOverlappingByMinPoint */
static int oc_overlap_minpoint(const double *mnL, const double *szL, const double *mnR, const double *szR, int D, double tol) {
  for (int d = 0; d < D; ++d) {
    double maxL = mnL[d] + szL[d];
    double maxR = mnR[d] + szR[d];
    if (maxL <= mnR[d] - tol || maxR <= mnL[d] - tol) return 0;
  }
  return 1;
}

/* internal_geometry_module.h:190-198 DoesRangeContainBox */
static inline int oc_range_contains(double rmin, double rmax, double bmin, double bmax, double tol) {
  if (rmin - tol > bmin || bmin > rmax + tol) return 0;
  if (rmin - tol > bmax || bmax > rmax + tol) return 0;
  return 1;
}

/* adapters/general.h:262-289 GetBoxRelation == Overlapped */
static int oc_boxes_overlap_strict(const double *e1, const double *e2, int D, double tol) {
  int adjacent = 0;
  for (int d = 0; d < D; ++d) {
    double min1 = e1[d], max1 = e1[D + d], min2 = e2[d], max2 = e2[D + d];
    if (min1 + tol < max2 && min2 + tol < max1) continue;
    if (fabs(min1 - max2) <= tol || fabs(max1 - min2) <= tol) {
      adjacent = 1;
      continue;
    }
    if (max2 + tol < min1 || max1 + tol < min2) return 0;
  }
  return !adjacent;
}

/* ------------------------------------------------------------------------- */
/* Result buffers                                                             */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint32_t *ids;
  size_t n, cap;
} oc_vec;

static void vec_push(oc_vec *v, uint32_t x) {
  if (v->n == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 16;
    v->ids = (uint32_t *)realloc(v->ids, v->cap * 4u);
  }
  v->ids[v->n++] = x;
}

typedef struct {
  uint64_t nq;
  uint64_t *offsets;
  uint32_t *ids;
  uint64_t nids;
} oc_result;

OC_API uint64_t oc_result_nids(const oc_result *r) { return r->nids; }
OC_API uint64_t oc_result_nq(const oc_result *r) { return r->nq; }
OC_API void oc_result_copy(const oc_result *r, uint64_t *offsets, uint32_t *ids) {
  if (offsets) memcpy(offsets, r->offsets, 8u * (r->nq + 1));
  if (ids && r->nids) memcpy(ids, r->ids, 4u * r->nids);
}
OC_API void oc_result_free(oc_result *r) {
  if (!r) return;
  free(r->offsets);
  free(r->ids);
  free(r);
}

/* ------------------------------------------------------------------------- */
/* kNN                                                                        */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint32_t id;
  double dist;
} oc_ed;

/* max-heap on distance (std::make_heap with operator<), ties broken by id so
 * the oracle is deterministic */
static inline int ed_less(const oc_ed *a, const oc_ed *b) {
  return a->dist < b->dist || (a->dist == b->dist && a->id < b->id);
}
static void heap_sift_up(oc_ed *h, size_t i) {
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (!ed_less(&h[p], &h[i])) break;
    oc_ed tmp = h[p];
    h[p] = h[i];
    h[i] = tmp;
    i = p;
  }
}
static void heap_sift_down(oc_ed *h, size_t n, size_t i) {
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < n && ed_less(&h[m], &h[l])) m = l;
    if (r < n && ed_less(&h[m], &h[r])) m = r;
    if (m == i) break;
    oc_ed tmp = h[m];
    h[m] = h[i];
    h[i] = tmp;
    i = m;
  }
}
static void heap_push(oc_ed *h, size_t *n, oc_ed x) {
  h[*n] = x;
  heap_sift_up(h, *n);
  ++*n;
}
static void heap_pop(oc_ed *h, size_t *n) {
  --*n;
  h[0] = h[*n];
  heap_sift_down(h, *n, 0);
}

typedef struct {
  uint32_t node;
  double prio;
} oc_pn;
/* min-heap of nodes by priority (std::priority_queue with std::greater) */
static void pq_push(oc_pn **h, size_t *n, size_t *cap, oc_pn x) {
  if (*n == *cap) {
    *cap = *cap ? *cap * 2 : 64;
    *h = (oc_pn *)realloc(*h, *cap * sizeof(oc_pn));
  }
  size_t i = (*n)++;
  (*h)[i] = x;
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (!((*h)[i].prio < (*h)[p].prio)) break;
    oc_pn t = (*h)[p];
    (*h)[p] = (*h)[i];
    (*h)[i] = t;
    i = p;
  }
}
static oc_pn pq_pop(oc_pn *h, size_t *n) {
  oc_pn top = h[0];
  --*n;
  h[0] = h[*n];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < *n && h[l].prio < h[m].prio) m = l;
    if (r < *n && h[r].prio < h[m].prio) m = r;
    if (m == i) break;
    oc_pn t = h[m];
    h[m] = h[i];
    h[i] = t;
    i = m;
  }
  return top;
}

static int ed_cmp_asc(const void *a, const void *b) {
  const oc_ed *x = (const oc_ed *)a, *y = (const oc_ed *)b;
  if (x->dist != y->dist) return x->dist < y->dist ? -1 : 1;
  return x->id < y->id ? -1 : (x->id > y->id);
}

/* ot_query.h:1033-1036 GetValueWithToleranceUpper */
static inline double tol_upper(double v, double tol) { return fmax(tol, v * (1.0 + tol)); }

/* adapters/general.h:159-210 MinMaxDistance2 + :212-218 MinMaxDistance
 * (Roussopoulos MINMAXDIST): optimistic distance (0 inside the box, within
 * tol per axis) and pessimistic distance of a point to a box, same op order. */
static void oc_minmax_distance(const double *pt, const double *box, int D, double tol, double *dmin,
                               double *dminmax) {
  double min2 = 0.0, minmax2 = 0.0;
  double farthest_inside2 = DBL_MAX, largest_diff = 0.0;
  int inside = 1;
  for (int d = 0; d < D; ++d) {
    const double v = pt[d], minC = box[d], maxC = box[D + d];
    const double minDist = v - minC, maxDist = maxC - v;
    const int in_c = (-tol <= minDist && -tol <= maxDist);
    inside &= in_c;
    double a = minDist * minDist, b = maxDist * maxDist;
    if (b < a) {
      double tmp = a;
      a = b;
      b = tmp;
    }
    if (inside) farthest_inside2 = (b < farthest_inside2) ? b : farthest_inside2; /* std::min */
    if (!in_c) min2 += a;
    largest_diff = (largest_diff < b - a) ? b - a : largest_diff; /* std::max */
    minmax2 += b;
  }
  if (inside) {
    min2 = 0.0;
    minmax2 = farthest_inside2;
  } else {
    minmax2 -= largest_diff;
  }
  *dmin = sqrt(min2);
  *dminmax = sqrt(minmax2);
}

/* ot_query.h:1208-1247 GetNearestNeighbors with AddEntityDistance :1058-1164
 * (optimistic heap by the optimistic distance, pessimistic heap of the k
 * smallest pessimistic distances, F from the pessimistic k-th, eviction of
 * optimistic entries > F together with their pessimistic entries) and the
 * priority traversal :235-291 on GetBoxWallDistanceAD(.., true).  Points use
 * Distance for both; boxes MinMaxDistance (:1079-1082).  Writes at most out_cap
 * ids; returns the full result size. */
static size_t oc_knn_one(const oc_tree *t, const double *pts, const double *q, uint32_t k, double max_within, double tol,
                         uint32_t *out, size_t out_cap, oc_pn **pqbuf, size_t *pqcap, oc_ed **ebuf, size_t *ecap) {
  int D = t->dim;
  if (k == 0 || t->node_count == 0) return 0;
  double F = (max_within == DBL_MAX) ? DBL_MAX : tol_upper(max_within, tol);
  size_t need = (size_t)t->n + k + 4;
  if (*ecap < 2 * need) {
    *ecap = 2 * need;
    *ebuf = (oc_ed *)realloc(*ebuf, *ecap * sizeof(oc_ed));
  }
  oc_ed *opt = *ebuf;        /* optimistic max-heap (may exceed k) */
  oc_ed *pes = *ebuf + need; /* pessimistic max-heap (k) */
  size_t nopt = 0, npes = 0;
  int heapified = 0;
  size_t npq = 0;
  double w0 = oc_wall_distance(q, oc_nmin(t, 0), oc_nsize(t, 0), D, 1);
  if (!(w0 >= F)) pq_push(pqbuf, &npq, pqcap, (oc_pn){0, w0});
  while (npq) {
    oc_pn top = pq_pop(*pqbuf, &npq);
    if (top.prio >= F) break; /* Terminate */
    uint32_t node = top.node;
    for (uint32_t j = t->node_ebegin[node]; j < t->node_ebegin[node] + t->node_ecount[node]; ++j) {
      uint32_t id = t->entity_ids[j];
      double dmin, dminmax;
      if (t->is_box) {
        oc_minmax_distance(q, &pts[(size_t)id * 2 * D], D, tol, &dmin, &dminmax);
      } else {
        dmin = dminmax = oc_distance(q, &pts[(size_t)id * D], D);
      }
      if (dmin >= F) continue;
      oc_ed e = {id, dmin}, ep = {id, dminmax};
      if (!heapified) {
        opt[nopt++] = e;
        pes[npes++] = ep;
        if (npes < k) continue;
        /* the optimistic size was k-1 before this push: make_heap both, F from the k-th */
        for (size_t i = 0; i < nopt; ++i) heap_sift_up(opt, i);
        for (size_t i = 0; i < npes; ++i) heap_sift_up(pes, i);
        heapified = 1;
        F = tol_upper(pes[0].dist, tol);
      } else {
        heap_push(opt, &nopt, e);
        heap_push(pes, &npes, ep);
        heap_pop(pes, &npes);
        double lim = tol_upper(pes[0].dist, tol);
        if (lim < F) {
          F = lim;
          int pes_removed = 0;
          while (nopt && opt[0].dist > F) {
            uint32_t rid = opt[0].id;
            for (size_t i = 0; i < npes; ++i)
              if (pes[i].id == rid) {
                pes[i] = pes[--npes];
                pes_removed = 1;
                break;
              }
            heap_pop(opt, &nopt);
          }
          if (pes_removed)
            for (size_t i = 0; i < npes; ++i) heap_sift_up(pes, i);
        }
      }
    }
    for (uint32_t c = t->node_cbegin[node]; c < t->node_cbegin[node] + t->node_ccount[node]; ++c) {
      double w = oc_wall_distance(q, oc_nmin(t, c), oc_nsize(t, c), D, 1);
      if (w >= F) continue;
      pq_push(pqbuf, &npq, pqcap, (oc_pn){c, w});
    }
  }
  qsort(opt, nopt, sizeof(oc_ed), ed_cmp_asc);
  for (size_t i = 0; i < nopt && i < out_cap; ++i) out[i] = opt[i].id;
  return nopt;
}

/* ------------------------------------------------------------------------- */
/* Range search                                                               */
/* ------------------------------------------------------------------------- */

/* ot_query.h:476-583 RangeSearch: root shortcut -> iota, zero volume -> empty,
 * DFS (last child popped first) with the contained-subtree shortcut.  Points:
 * DoesBoxContainPoint (general.h:225-243); boxes: Inside = AreBoxesOverlapped
 * (range, e, true) (general.h:296-308), Overlap = GetBoxRelation == Overlapped.
 * Output order is the reference's DFS order. */
static void oc_range_one(const oc_tree *t, const double *geom, const double *range, int overlap_mode, double tol, oc_vec *out,
                         uint32_t **stack, size_t *scap) {
  int D = t->dim;
  if (t->node_count == 0) return;
  /* root box: GetNodeBox(root) */
  const double *rmn = oc_nmin(t, 0), *rsz = oc_nsize(t, 0);
  int contains_root = 1;
  for (int d = 0; d < D; ++d) {
    double bmax = rmn[d] + rsz[d];
    if (!oc_range_contains(range[d], range[D + d], rmn[d], bmax, tol)) {
      contains_root = 0;
      break;
    }
  }
  if (contains_root) {
    for (uint32_t i = 0; i < t->n; ++i) vec_push(out, i);
    return;
  }
  double vol = 1.0;
  for (int d = 0; d < D; ++d) vol *= range[D + d] - range[d];
  if (vol <= 0.0) return;
  double smn[OC_MAXD], ssz[OC_MAXD];
  for (int d = 0; d < D; ++d) {
    smn[d] = range[d];
    ssz[d] = range[D + d] - range[d];
  }
  size_t sp = 0;
  if (*scap < 64) {
    *scap = 64;
    *stack = (uint32_t *)realloc(*stack, *scap * 4u);
  }
  (*stack)[sp++] = 0;
  while (sp) {
    uint32_t node = (*stack)[--sp];
    const double *nmn = oc_nmin(t, node), *nsz = oc_nsize(t, node);
    if (!oc_overlap_minpoint(smn, ssz, nmn, nsz, D, tol)) continue;
    int contained = 1;
    for (int d = 0; d < D; ++d) {
      double bmax = nmn[d] + nsz[d];
      if (!oc_range_contains(range[d], range[D + d], nmn[d], bmax, tol)) {
        contained = 0;
        break;
      }
    }
    if (contained) {
      /* nested DFS over the subtree, copying every node's entities */
      size_t base = sp;
      (*stack)[sp++] = node;
      while (sp > base) {
        uint32_t s = (*stack)[--sp];
        for (uint32_t j = t->node_ebegin[s]; j < t->node_ebegin[s] + t->node_ecount[s]; ++j) vec_push(out, t->entity_ids[j]);
        for (uint32_t c = t->node_cbegin[s]; c < t->node_cbegin[s] + t->node_ccount[s]; ++c) {
          if (sp == *scap) {
            *scap *= 2;
            *stack = (uint32_t *)realloc(*stack, *scap * 4u);
          }
          (*stack)[sp++] = c;
        }
      }
      continue;
    }
    for (uint32_t j = t->node_ebegin[node]; j < t->node_ebegin[node] + t->node_ecount[node]; ++j) {
      uint32_t id = t->entity_ids[j];
      int ok;
      if (!t->is_box) {
        const double *p = &geom[(size_t)id * D];
        ok = 1;
        if (tol != 0.0) {
          for (int d = 0; d < D; ++d)
            if (!(range[d] - tol < p[d] && p[d] < range[D + d] + tol)) {
              ok = 0;
              break;
            }
        } else {
          for (int d = 0; d < D; ++d)
            if (!(range[d] <= p[d] && p[d] <= range[D + d])) {
              ok = 0;
              break;
            }
        }
      } else {
        const double *e = &geom[(size_t)id * 2 * D];
        if (overlap_mode) ok = oc_boxes_overlap_strict(range, e, D, tol);
        else {
          ok = 1;
          for (int d = 0; d < D; ++d) {
            if (range[d] > e[d] || e[d] > range[D + d]) {
              ok = 0;
              break;
            }
            if (range[d] > e[D + d] || e[D + d] > range[D + d]) {
              ok = 0;
              break;
            }
          }
        }
      }
      if (ok) vec_push(out, id);
    }
    for (uint32_t c = t->node_cbegin[node]; c < t->node_cbegin[node] + t->node_ccount[node]; ++c) {
      if (sp == *scap) {
        *scap *= 2;
        *stack = (uint32_t *)realloc(*stack, *scap * 4u);
      }
      (*stack)[sp++] = c;
    }
  }
}

/* internal_geometry_module.h:372-396 DoesBoxContainPointAD(box, point, tol) */
static int oc_box_contains_point(const double *mn, const double *mx, const double *p, int D, double tol) {
  if (tol != 0.0) {
    for (int d = 0; d < D; ++d)
      if (!(mn[d] - tol < p[d] && p[d] < mx[d] + tol)) return 0;
  } else {
    for (int d = 0; d < D; ++d)
      if (!(mn[d] <= p[d] && p[d] <= mx[d])) return 0;
  }
  return 1;
}

/* internal_geometry_module.h:398-422 DoesBoxContainPointAD(minPoint, size, point, tol):
 * centre form |p - min - size/2| against size/2 (+ tol). */
static int oc_node_contains_point(const double *mn, const double *sz, const double *p, int D, double tol) {
  for (int d = 0; d < D; ++d) {
    double half = sz[d] * 0.5;
    double dist = fabs(p[d] - mn[d] - half);
    if (tol != 0.0) {
      if (dist >= half + tol) return 0;
    } else if (dist > half) {
      return 0;
    }
  }
  return 1;
}

/* ot_query.h:426-466 PickSearch: root box test (GetNodeBox, min/max form), then
 * DFS (last child popped first) skipping nodes whose box does not contain the
 * pick point; entities: points ArePointsEqual (general.h:218-221, Distance2 <=
 * tol^2), boxes DoesBoxContainPoint (general.h:225-243).  Output in DFS order. */
static void oc_pick_one(const oc_tree *t, const double *geom, const double *p, double tol, oc_vec *out, uint32_t **stack,
                        size_t *scap) {
  int D = t->dim;
  if (t->node_count == 0) return;
  double rmx[OC_MAXD];
  for (int d = 0; d < D; ++d) rmx[d] = oc_nmin(t, 0)[d] + oc_nsize(t, 0)[d];
  if (!oc_box_contains_point(oc_nmin(t, 0), rmx, p, D, tol)) return;
  size_t sp = 0;
  if (*scap < 64) {
    *scap = 64;
    *stack = (uint32_t *)realloc(*stack, *scap * 4u);
  }
  (*stack)[sp++] = 0;
  while (sp) {
    uint32_t node = (*stack)[--sp];
    if (!oc_node_contains_point(oc_nmin(t, node), oc_nsize(t, node), p, D, tol)) continue;
    for (uint32_t j = t->node_ebegin[node]; j < t->node_ebegin[node] + t->node_ecount[node]; ++j) {
      uint32_t id = t->entity_ids[j];
      int ok;
      if (t->is_box) {
        const double *e = &geom[(size_t)id * 2 * D];
        ok = oc_box_contains_point(e, e + D, p, D, tol);
      } else {
        const double *e = &geom[(size_t)id * D];
        double d2 = 0.0;
        for (int d = 0; d < D; ++d) {
          double x = e[d] - p[d];
          d2 += x * x;
        }
        ok = d2 <= tol * tol;
      }
      if (ok) vec_push(out, id);
    }
    for (uint32_t c = t->node_cbegin[node]; c < t->node_cbegin[node] + t->node_ccount[node]; ++c) {
      if (sp == *scap) {
        *scap *= 2;
        *stack = (uint32_t *)realloc(*stack, *scap * 4u);
      }
      (*stack)[sp++] = c;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Planes: PlaneSearch, PlanePositiveSegmentation, FrustumCulling             */
/* ------------------------------------------------------------------------- */
enum { OC_REL_NEGATIVE = -1, OC_REL_HIT = 0, OC_REL_POSITIVE = 1 };

/* internal_geometry_module.h:254-280 GetBoxPlaneRelationAD(minPoint, size, d, n, tol) */
static int oc_box_plane_relation(const double *mn, const double *sz, double dist, const double *n, int D, double tol) {
  double center[OC_MAXD], half[OC_MAXD];
  for (int d = 0; d < D; ++d) {
    half[d] = sz[d] * 0.5;
    center[d] = mn[d] + half[d];
  }
  double radius = tol;
  for (int d = 0; d < D; ++d) radius += half[d] * fabs(n[d]);
  double dot = 0.0; /* DotAD(planeNormal, center) :180-188 */
  for (int d = 0; d < D; ++d) dot += n[d] * center[d];
  double cp = dot - dist;
  if (cp + radius < 0.0) return OC_REL_NEGATIVE;
  if (cp - radius > 0.0) return OC_REL_POSITIVE;
  return OC_REL_HIT;
}

/* ot_query.h:585-602 GetEntityPlaneRelation: points GetPointPlaneRelation
 * (general.h:322-335), boxes GetBoxPlaneRelationAD(min, max - min). */
static int oc_entity_plane_relation(const oc_tree *t, const double *geom, uint32_t id, double dist, const double *n, double tol) {
  int D = t->dim;
  if (t->is_box) {
    const double *e = &geom[(size_t)id * 2 * D];
    double sz[OC_MAXD];
    for (int d = 0; d < D; ++d) sz[d] = e[D + d] - e[d];
    return oc_box_plane_relation(e, sz, dist, n, D, tol);
  }
  const double *p = &geom[(size_t)id * D];
  double proj = 0.0;
  for (int d = 0; d < D; ++d) proj += n[d] * p[d];
  if (proj < dist - tol) return OC_REL_NEGATIVE;
  if (proj > dist + tol) return OC_REL_POSITIVE;
  return OC_REL_HIT;
}

enum { OC_PLANE_SEARCH = 0, OC_PLANE_POSITIVE = 1, OC_FRUSTUM = 2 };

/* One query over planes[nplanes] (each: origo distance, then D normal components).
 * OC_PLANE_SEARCH ot_query.h:612-636: DFS into nodes the plane hits, entities that
 * the plane hits.  OC_PLANE_POSITIVE :691-741: nodes on the negative side are
 * skipped, a node wholly on the positive side contributes its whole subtree (no
 * entity test), otherwise entities not on the negative side.  OC_FRUSTUM :756-822:
 * a node is visited unless a plane, met before any plane that hits it, has it on
 * the negative side; an entity is kept when no plane has it on the negative side. */
static void oc_planes_one(const oc_tree *t, const double *geom, const double *planes, uint32_t nplanes, int mode, double tol,
                          oc_vec *out, uint32_t **stack, size_t *scap) {
  int D = t->dim;
  if (t->node_count == 0 || nplanes == 0) return;
  size_t sp = 0;
  if (*scap < 64) {
    *scap = 64;
    *stack = (uint32_t *)realloc(*stack, *scap * 4u);
  }
  (*stack)[sp++] = 0;
  while (sp) {
    uint32_t node = (*stack)[--sp];
    const double *mn = oc_nmin(t, node), *sz = oc_nsize(t, node);
    int visit = 1, whole = 0;
    if (mode == OC_FRUSTUM) {
      for (uint32_t k = 0; k < nplanes; ++k) {
        const double *pl = &planes[(size_t)k * (D + 1)];
        int rel = oc_box_plane_relation(mn, sz, pl[0], pl + 1, D, tol);
        if (rel == OC_REL_HIT) break;
        if (rel == OC_REL_NEGATIVE) {
          visit = 0;
          break;
        }
      }
    } else {
      int rel = oc_box_plane_relation(mn, sz, planes[0], planes + 1, D, tol);
      if (mode == OC_PLANE_SEARCH) visit = rel == OC_REL_HIT;
      else if (rel == OC_REL_NEGATIVE) visit = 0;
      else if (rel == OC_REL_POSITIVE) whole = 1;
    }
    if (!visit) continue;
    if (whole) {
      /* nested DFS over the subtree, every node's entities */
      size_t base = sp;
      (*stack)[sp++] = node;
      while (sp > base) {
        uint32_t s = (*stack)[--sp];
        for (uint32_t j = t->node_ebegin[s]; j < t->node_ebegin[s] + t->node_ecount[s]; ++j) vec_push(out, t->entity_ids[j]);
        for (uint32_t c = t->node_cbegin[s]; c < t->node_cbegin[s] + t->node_ccount[s]; ++c) {
          if (sp == *scap) {
            *scap *= 2;
            *stack = (uint32_t *)realloc(*stack, *scap * 4u);
          }
          (*stack)[sp++] = c;
        }
      }
      continue;
    }
    for (uint32_t j = t->node_ebegin[node]; j < t->node_ebegin[node] + t->node_ecount[node]; ++j) {
      uint32_t id = t->entity_ids[j];
      int ok;
      if (mode == OC_PLANE_SEARCH) {
        ok = oc_entity_plane_relation(t, geom, id, planes[0], planes + 1, tol) == OC_REL_HIT;
      } else {
        ok = 1;
        for (uint32_t k = 0; k < nplanes && ok; ++k) {
          const double *pl = &planes[(size_t)k * (D + 1)];
          ok = oc_entity_plane_relation(t, geom, id, pl[0], pl + 1, tol) != OC_REL_NEGATIVE;
        }
      }
      if (ok) vec_push(out, id);
    }
    for (uint32_t c = t->node_cbegin[node]; c < t->node_cbegin[node] + t->node_ccount[node]; ++c) {
      if (sp == *scap) {
        *scap *= 2;
        *stack = (uint32_t *)realloc(*stack, *scap * 4u);
      }
      (*stack)[sp++] = c;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Ray first hit                                                              */
/* ------------------------------------------------------------------------- */
typedef struct {
  double origin[OC_MAXD], inv[OC_MAXD], dir[OC_MAXD];
  double min_tol, inc;
  int sign[OC_MAXD], zero[OC_MAXD];
  int has_nan;
} oc_ray;

/* internal_geometry_module.h:539-568 RayHitTester::Make */
static int oc_ray_make(oc_ray *r, const double *o, const double *dir, int D, double tol, double inc) {
  double s2 = 0.0;
  for (int d = 0; d < D; ++d) s2 += dir[d] * dir[d];
  if (!(fabs(s2 - 1.0) < 0.000001)) return 0; /* general.h:223 IsNormalizedVector */
  r->min_tol = tol;
  r->inc = inc;
  r->has_nan = 0;
  for (int d = 0; d < D; ++d) {
    r->origin[d] = o[d];
    r->dir[d] = dir[d];
    int z = fabs(dir[d]) < 2.0 * DBL_EPSILON;
    r->inv[d] = z ? NAN : (1.0 / dir[d]);
    r->sign[d] = dir[d] < 0;
    r->zero[d] = z;
    r->has_nan |= z;
  }
  return 1;
}

/* internal_geometry_module.h:631-712 BoxHitTest<isConeToleranceConsidered=true> */
static int oc_box_hit(const oc_ray *r, const double *mind, const double *maxd, int D, double *enter, double *exit) {
  double en = -DBL_MAX, ex = DBL_MAX;
  double pd0[OC_MAXD], pd1[OC_MAXD];
  for (int d = 0; d < D; ++d) pd0[d] = mind[d] * r->inv[d];
  for (int d = 0; d < D; ++d) pd1[d] = maxd[d] * r->inv[d];
  if (r->has_nan) {
    for (int d = 0; d < D; ++d)
      if (!r->zero[d]) en = fmax(en, r->sign[d] ? pd1[d] : pd0[d]);
    for (int d = 0; d < D; ++d)
      if (!r->zero[d]) ex = fmin(ex, r->sign[d] ? pd0[d] : pd1[d]);
  } else {
    for (int d = 0; d < D; ++d) en = smax(en, r->sign[d] ? pd1[d] : pd0[d]);
    for (int d = 0; d < D; ++d) ex = smin(ex, r->sign[d] ? pd0[d] : pd1[d]);
  }
  double exit_tol = smax(0.0, ex) * r->inc + r->min_tol;
  en -= 0.5 * (smax(0.0, en) * r->inc + r->min_tol);
  ex += 0.5 * (smax(0.0, ex) * r->inc + r->min_tol);
  if (en >= ex || ex <= 0.0) return 0;
  en = smax(0.0, en);
  if (r->has_nan) {
    for (int d = 0; d < D; ++d) {
      if (!r->zero[d]) continue;
      if (mind[d] > exit_tol || maxd[d] < -exit_tol) return 0;
    }
  }
  *enter = en;
  *exit = ex;
  return 1;
}

/* Hit(minPoint, size) :575-585 */
static int oc_hit_node(const oc_ray *r, const double *mn, const double *sz, int D, double *en, double *ex) {
  double a[OC_MAXD], b[OC_MAXD];
  for (int d = 0; d < D; ++d) a[d] = mn[d] - r->origin[d];
  for (int d = 0; d < D; ++d) b[d] = mn[d] + sz[d] - r->origin[d];
  return oc_box_hit(r, a, b, D, en, ex);
}
/* Hit(box) :587-597 */
static int oc_hit_box(const oc_ray *r, const double *bx, int D, double *en, double *ex) {
  double a[OC_MAXD], b[OC_MAXD];
  for (int d = 0; d < D; ++d) a[d] = bx[d] - r->origin[d];
  for (int d = 0; d < D; ++d) b[d] = bx[D + d] - r->origin[d];
  return oc_box_hit(r, a, b, D, en, ex);
}
/* Hit(point) :599-626 */
static int oc_hit_point(const oc_ray *r, const double *p, int D, double *en, double *ex) {
  double v[OC_MAXD];
  for (int d = 0; d < D; ++d) v[d] = p[d] - r->origin[d];
  double tt = 0.0;
  for (int d = 0; d < D; ++d) tt += v[d] * r->dir[d];
  double len2 = 0.0;
  for (int d = 0; d < D; ++d) len2 += v[d] * v[d];
  double dist2 = smax(0.0, len2 - tt * tt);
  double ct = smax(0.0, tt) * r->inc + r->min_tol;
  if (dist2 <= ct * ct && tt >= -ct) {
    *en = smax(0.0, tt - ct);
    *ex = tt + ct;
    return 1;
  }
  return 0;
}

typedef struct {
  uint32_t id;
  double enter;
} oc_cand;
static int cand_cmp_desc(const void *a, const void *b) {
  const oc_cand *x = (const oc_cand *)a, *y = (const oc_cand *)b;
  if (x->enter != y->enter) return x->enter > y->enter ? -1 : 1;
  return x->id < y->id ? -1 : (x->id > y->id);
}

/* ot_query.h:2035-2122 RayIntersectedFirst: priority traversal on node enter
 * distance, terminate when enter > M; candidates pushed with their enter
 * distance; M = (1+inc)*exit + appliedTol when exit < M, popping candidates
 * with enter > M.  Output = heap pops = descending enter (ties by id). */
static size_t oc_ray_one(const oc_tree *t, const double *geom, const double *o, const double *dir, double tol, double inc,
                         double max_dist, oc_vec *out, oc_pn **pqbuf, size_t *pqcap) {
  int D = t->dim;
  oc_ray r;
  if (t->node_count == 0 || !oc_ray_make(&r, o, dir, D, tol, inc)) return 0;
  double applied = (tol == 0) ? DBL_EPSILON : tol;
  double M = max_dist;
  size_t ncand = 0, ccap = 16;
  oc_cand *cand = (oc_cand *)malloc(ccap * sizeof(oc_cand));
  size_t npq = 0;
  double en, ex;
  if (oc_hit_node(&r, oc_nmin(t, 0), oc_nsize(t, 0), D, &en, &ex)) pq_push(pqbuf, &npq, pqcap, (oc_pn){0, en});
  while (npq) {
    oc_pn top = pq_pop(*pqbuf, &npq);
    if (top.prio > M) break;
    uint32_t node = top.node;
    for (uint32_t j = t->node_ebegin[node]; j < t->node_ebegin[node] + t->node_ecount[node]; ++j) {
      uint32_t id = t->entity_ids[j];
      int hit = t->is_box ? oc_hit_box(&r, &geom[(size_t)id * 2 * D], D, &en, &ex) : oc_hit_point(&r, &geom[(size_t)id * D], D, &en, &ex);
      if (!hit) continue;
      if (en > M) continue;
      if (ncand == ccap) {
        ccap *= 2;
        cand = (oc_cand *)realloc(cand, ccap * sizeof(oc_cand));
      }
      cand[ncand++] = (oc_cand){id, en};
      if (ex >= M) continue;
      M = (1.0 + inc) * ex + applied;
      size_t w = 0;
      for (size_t i = 0; i < ncand; ++i)
        if (!(cand[i].enter > M)) cand[w++] = cand[i];
      ncand = w;
    }
    for (uint32_t c = t->node_cbegin[node]; c < t->node_cbegin[node] + t->node_ccount[node]; ++c)
      if (oc_hit_node(&r, oc_nmin(t, c), oc_nsize(t, c), D, &en, &ex)) pq_push(pqbuf, &npq, pqcap, (oc_pn){c, en});
  }
  qsort(cand, ncand, sizeof(oc_cand), cand_cmp_desc);
  for (size_t i = 0; i < ncand; ++i) vec_push(out, cand[i].id);
  free(cand);
  return ncand;
}

static int cand_cmp_asc(const void *a, const void *b) {
  const oc_cand *x = (const oc_cand *)a, *y = (const oc_cand *)b;
  if (x->enter != y->enter) return x->enter < y->enter ? -1 : 1;
  return x->id < y->id ? -1 : (x->id > y->id);
}

/* ot_query.h:1929-2012 RayIntersectedAll<SHOULD_SORT = true>: DFS skipping nodes
 * the ray misses (Hit(minPoint, size)); every entity hit whose enter distance is
 * within maxExaminationDistance (or any, when it is 0) is kept with its enter
 * distance; the list is sorted ascending by enter distance (ties by id here: the
 * reference's std::ranges::sort leaves them unspecified). */
static size_t oc_ray_all_one(const oc_tree *t, const double *geom, const double *o, const double *dir, double tol, double inc,
                             double max_dist, oc_vec *out, uint32_t **stack, size_t *scap) {
  int D = t->dim;
  oc_ray r;
  if (t->node_count == 0 || !oc_ray_make(&r, o, dir, D, tol, inc)) return 0;
  size_t ncand = 0, ccap = 16;
  oc_cand *cand = (oc_cand *)malloc(ccap * sizeof(oc_cand));
  size_t sp = 0;
  if (*scap < 64) {
    *scap = 64;
    *stack = (uint32_t *)realloc(*stack, *scap * 4u);
  }
  (*stack)[sp++] = 0;
  double en, ex;
  while (sp) {
    uint32_t node = (*stack)[--sp];
    if (!oc_hit_node(&r, oc_nmin(t, node), oc_nsize(t, node), D, &en, &ex)) continue;
    for (uint32_t j = t->node_ebegin[node]; j < t->node_ebegin[node] + t->node_ecount[node]; ++j) {
      uint32_t id = t->entity_ids[j];
      int hit = t->is_box ? oc_hit_box(&r, &geom[(size_t)id * 2 * D], D, &en, &ex) : oc_hit_point(&r, &geom[(size_t)id * D], D, &en, &ex);
      if (!hit) continue;
      if (!(max_dist == 0 || en <= max_dist)) continue;
      if (ncand == ccap) {
        ccap *= 2;
        cand = (oc_cand *)realloc(cand, ccap * sizeof(oc_cand));
      }
      cand[ncand++] = (oc_cand){id, en};
    }
    for (uint32_t c = t->node_cbegin[node]; c < t->node_cbegin[node] + t->node_ccount[node]; ++c) {
      if (sp == *scap) {
        *scap *= 2;
        *stack = (uint32_t *)realloc(*stack, *scap * 4u);
      }
      (*stack)[sp++] = c;
    }
  }
  qsort(cand, ncand, sizeof(oc_cand), cand_cmp_asc);
  for (size_t i = 0; i < ncand; ++i) vec_push(out, cand[i].id);
  free(cand);
  return ncand;
}

/* ------------------------------------------------------------------------- */
/* Collision detection (boxes, self)                                          */
/* ------------------------------------------------------------------------- */
typedef struct {
  oc_vec pairs; /* flattened (a, b) */
} oc_pairs;

static void pairs_push(oc_pairs *p, uint32_t a, uint32_t b) {
  vec_push(&p->pairs, a);
  vec_push(&p->pairs, b);
}

/* per thread: collisions of different trees may run on concurrent threads */
static _Thread_local const double *g_sort_geom;
static _Thread_local int g_sort_dim;
/* ot_query.h:1699-1716 comparator: (minX, id) */
static int sort_minx_id(const void *a, const void *b) {
  uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
  double x1 = g_sort_geom[(size_t)i * 2 * g_sort_dim], x2 = g_sort_geom[(size_t)j * 2 * g_sort_dim];
  if (x1 < x2) return -1;
  if (x2 < x1) return 1;
  return i < j ? -1 : (i > j);
}
/* SweepAndPruneDatabase :1275-1300: sort by minX only (ties: id here) */
static uint32_t *sorted_node_entities(const oc_tree *t, const double *geom, uint32_t node, uint32_t *cnt) {
  uint32_t n = t->node_ecount[node];
  uint32_t *v = (uint32_t *)malloc(4u * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) v[i] = t->entity_ids[t->node_ebegin[node] + i];
  g_sort_geom = geom;
  g_sort_dim = t->dim;
  qsort(v, n, 4, sort_minx_id);
  *cnt = n;
  return v;
}

/* ot_query.h:1307-1458 two-tree CollisionDetection: breadth-first over node pairs
 * (left subtree x right subtree), sweep and prune on minX within each pair, a
 * traversed node kept as its own "child" so its entities meet the other side's
 * children.  Used for sibling subtrees of one tree (tl == tr) and for two trees. */
static void oc_collide_dual(const oc_tree *tl, const double *gl, uint32_t left, const oc_tree *tr, const double *gr, uint32_t right,
                            double tol, oc_pairs *out, uint32_t **cache_l, uint32_t *cache_ln, uint32_t **cache_r,
                            uint32_t *cache_rn) {
  int D = tl->dim;
  typedef struct {
    uint32_t l, r;
    uint8_t lt, rt;
  } pr;
  size_t qh = 0, qt = 0, qcap = 64;
  pr *q = (pr *)malloc(qcap * sizeof(pr));
  q[qt++] = (pr){left, right, 0, 0};
  while (qh < qt) {
    pr par = q[qh++];
    uint32_t *L, *R, nl, nr;
    if (!cache_l[par.l]) cache_l[par.l] = sorted_node_entities(tl, gl, par.l, &cache_ln[par.l]);
    if (!cache_r[par.r]) cache_r[par.r] = sorted_node_entities(tr, gr, par.r, &cache_rn[par.r]);
    L = cache_l[par.l];
    nl = cache_ln[par.l];
    R = cache_r[par.r];
    nr = cache_rn[par.r];
    size_t rb = 0;
    for (uint32_t a = 0; a < nl; ++a) {
      const double *lg = &gl[(size_t)L[a] * 2 * D];
      for (; rb < nr; ++rb) {
        const double *rg = &gr[(size_t)R[rb] * 2 * D];
        if (rg[D + 0] >= lg[0]) break;
      }
      for (size_t b = rb; b < nr; ++b) {
        const double *rg = &gr[(size_t)R[b] * 2 * D];
        if (lg[D + 0] < rg[0]) break;
        if (oc_boxes_overlap_strict(lg, rg, D, tol)) pairs_push(out, L[a], R[b]);
      }
    }
    /* children */
    uint32_t lc[1 << 10], rc[1 << 10];
    uint8_t ltr[1 << 10], rtr[1 << 10];
    uint32_t nlc = 0, nrc = 0;
    if (!par.lt)
      for (uint32_t c = tl->node_cbegin[par.l]; c < tl->node_cbegin[par.l] + tl->node_ccount[par.l]; ++c) {
        lc[nlc] = c;
        ltr[nlc++] = 0;
      }
    if (!par.rt)
      for (uint32_t c = tr->node_cbegin[par.r]; c < tr->node_cbegin[par.r] + tr->node_ccount[par.r]; ++c) {
        rc[nrc] = c;
        rtr[nrc++] = 0;
      }
    if (nlc == 0 && nrc == 0) continue;
    if (tl->node_ecount[par.l] != 0) {
      lc[nlc] = par.l;
      ltr[nlc++] = 1;
    }
    if (tr->node_ecount[par.r] != 0) {
      rc[nrc] = par.r;
      rtr[nrc++] = 1;
    }
    for (uint32_t a = 0; a < nlc; ++a)
      for (uint32_t b = 0; b < nrc; ++b) {
        if (lc[a] == par.l && rc[b] == par.r) continue;
        if (!oc_overlap_minpoint(oc_nmin(tl, lc[a]), oc_nsize(tl, lc[a]), oc_nmin(tr, rc[b]), oc_nsize(tr, rc[b]), D, tol)) continue;
        if (qt == qcap) {
          qcap *= 2;
          q = (pr *)realloc(q, qcap * sizeof(pr));
        }
        q[qt++] = (pr){lc[a], rc[b], ltr[a], rtr[b]};
      }
  }
  free(q);
}

typedef struct {
  uint32_t node;
  uint32_t *ids;
  uint32_t n;
  double bmin[OC_MAXD], bmax[OC_MAXD];
} oc_ctx;

/* ot_query.h:1655-1687 InsertCollidedEntitiesInSubtree (SEQ path of
 * CollectCollidedEntities :1736-1746), with InsideNode :1518-1561,
 * WithParents :1563-1653 and the loose sibling pass :1669-1683. */
static void oc_collide_subtree(const oc_tree *t, const double *geom, uint32_t node, oc_ctx *stack, uint32_t depth, double tol,
                               oc_pairs *out, uint32_t **cache_ids, uint32_t *cache_n) {
  int D = t->dim;
  oc_ctx *ctx = &stack[depth];
  ctx->node = node;
  ctx->ids = sorted_node_entities(t, geom, node, &ctx->n);
  for (int d = 0; d < D; ++d) {
    ctx->bmin[d] = oc_nmin(t, node)[d];
    ctx->bmax[d] = oc_nmin(t, node)[d] + oc_nsize(t, node)[d];
  }
  /* inside node */
  for (uint32_t i = 0; i < ctx->n; ++i) {
    const double *bi = &geom[(size_t)ctx->ids[i] * 2 * D];
    for (uint32_t j = i + 1; j < ctx->n; ++j) {
      const double *bj = &geom[(size_t)ctx->ids[j] * 2 * D];
      if (bi[D + 0] < bj[0]) break;
      if (oc_boxes_overlap_strict(bi, bj, D, tol)) pairs_push(out, ctx->ids[i], ctx->ids[j]);
    }
  }
  /* with parents */
  const double *nsz = oc_nsize(t, node);
  for (uint32_t pd = 0; pd < depth; ++pd) {
    oc_ctx *par = &stack[pd];
    size_t ib = 0;
    for (uint32_t a = 0; a < par->n; ++a) {
      const double *pg = &geom[(size_t)par->ids[a] * 2 * D];
      if (pg[0] > ctx->bmax[0]) break;
      double pmn[OC_MAXD], psz[OC_MAXD];
      for (int d = 0; d < D; ++d) {
        pmn[d] = pg[d];
        psz[d] = pg[D + d] - pg[d];
      }
      if (!oc_overlap_minpoint(ctx->bmin, nsz, pmn, psz, D, tol)) continue;
      for (; ib < ctx->n; ++ib)
        if (geom[(size_t)ctx->ids[ib] * 2 * D + D] >= pg[0]) break;
      for (size_t e = ib; e < ctx->n; ++e) {
        const double *eg = &geom[(size_t)ctx->ids[e] * 2 * D];
        if (pg[D + 0] < eg[0]) break;
        if (oc_boxes_overlap_strict(eg, pg, D, tol)) pairs_push(out, ctx->ids[e], par->ids[a]);
      }
    }
  }
  for (uint32_t c = t->node_cbegin[node]; c < t->node_cbegin[node] + t->node_ccount[node]; ++c)
    oc_collide_subtree(t, geom, c, stack, depth + 1, tol, out, cache_ids, cache_n);
  /* loose tree: sibling subtree pairs */
  for (uint32_t c1 = t->node_cbegin[node]; c1 < t->node_cbegin[node] + t->node_ccount[node]; ++c1)
    for (uint32_t c2 = c1 + 1; c2 < t->node_cbegin[node] + t->node_ccount[node]; ++c2)
      oc_collide_dual(t, geom, c1, t, geom, c2, tol, out, cache_ids, cache_n, cache_ids, cache_n);
  free(ctx->ids);
  ctx->ids = NULL;
}

static int pair_cmp(const void *a, const void *b) {
  const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : (x[1] > y[1]);
}

/* ot_query.h:1901-1906 CollisionDetection(entities, tol); pairs normalised to
 * (i < j) and sorted (the reference's orientation/order is unspecified). */
OC_API oc_result *oc_collision(const oc_tree *t, const double *geom, double tol) {
  oc_result *r = (oc_result *)calloc(1, sizeof(oc_result));
  oc_pairs p = {{0}};
  if (t->node_count && t->is_box) {
    oc_ctx stack[64];
    uint32_t **cache_ids = (uint32_t **)calloc(t->node_count, sizeof(uint32_t *));
    uint32_t *cache_n = (uint32_t *)calloc(t->node_count, 4);
    oc_collide_subtree(t, geom, 0, stack, 0, tol, &p, cache_ids, cache_n);
    for (uint32_t i = 0; i < t->node_count; ++i) free(cache_ids[i]);
    free(cache_ids);
    free(cache_n);
  }
  size_t np = p.pairs.n / 2;
  for (size_t i = 0; i < np; ++i) {
    uint32_t *pp = &p.pairs.ids[2 * i];
    if (pp[0] > pp[1]) {
      uint32_t tmp = pp[0];
      pp[0] = pp[1];
      pp[1] = tmp;
    }
  }
  if (np) qsort(p.pairs.ids, np, 8, pair_cmp);
  r->nq = np;
  r->nids = 2 * np;
  r->ids = p.pairs.ids;
  r->offsets = (uint64_t *)calloc(np + 1, 8);
  for (size_t i = 0; i <= np; ++i) r->offsets[i] = 2 * i;
  return r;
}

/* ot_query.h:1462-1484 static CollisionDetection(leftTree, leftEntities,
 * rightTree, rightEntities, tol): the dual walk from the two roots.  Pairs are
 * (left id, right id), sorted (the reference's order is its walk order). */
OC_API oc_result *oc_collision_trees(const oc_tree *tl, const double *gl, const oc_tree *tr, const double *gr, double tol) {
  oc_result *r = (oc_result *)calloc(1, sizeof(oc_result));
  oc_pairs p = {{0}};
  if (tl->node_count && tr->node_count && tl->is_box && tr->is_box && tl->dim == tr->dim) {
    uint32_t **cl = (uint32_t **)calloc(tl->node_count, sizeof(uint32_t *));
    uint32_t *cln = (uint32_t *)calloc(tl->node_count, 4);
    uint32_t **cr = (uint32_t **)calloc(tr->node_count, sizeof(uint32_t *));
    uint32_t *crn = (uint32_t *)calloc(tr->node_count, 4);
    oc_collide_dual(tl, gl, 0, tr, gr, 0, tol, &p, cl, cln, cr, crn);
    for (uint32_t i = 0; i < tl->node_count; ++i) free(cl[i]);
    for (uint32_t i = 0; i < tr->node_count; ++i) free(cr[i]);
    free(cl);
    free(cln);
    free(cr);
    free(crn);
  }
  size_t np = p.pairs.n / 2;
  if (np) qsort(p.pairs.ids, np, 8, pair_cmp);
  r->nq = np;
  r->nids = 2 * np;
  r->ids = p.pairs.ids;
  r->offsets = (uint64_t *)calloc(np + 1, 8);
  for (size_t i = 0; i <= np; ++i) r->offsets[i] = 2 * i;
  return r;
}

/* ------------------------------------------------------------------------- */
/* Batched query drivers (pthreads over contiguous query chunks)              */
/* ------------------------------------------------------------------------- */
enum { OC_Q_KNN = 0, OC_Q_RANGE = 1, OC_Q_RAY = 2, OC_Q_PICK = 3, OC_Q_RAY_ALL = 4, OC_Q_PLANES = 5 };

typedef struct {
  const oc_tree *t;
  const double *geom;
  const double *queries;
  int kind;
  uint32_t k;
  double max_within, tol, inc, max_dist;
  int overlap_mode;
  uint32_t nplanes; /* OC_Q_PLANES: planes per query; overlap_mode holds the plane mode */
  uint64_t q0, q1;
  oc_vec out;
  uint64_t *counts;
} oc_job;

static void *oc_job_run(void *arg) {
  oc_job *j = (oc_job *)arg;
  int D = j->t->dim;
  oc_pn *pq = NULL;
  size_t pqcap = 0;
  oc_ed *eb = NULL;
  size_t ecap = 0;
  uint32_t *stack = NULL;
  size_t scap = 0;
  uint32_t *tmp = NULL;
  size_t tcap = 0;
  for (uint64_t q = j->q0; q < j->q1; ++q) {
    size_t before = j->out.n;
    if (j->kind == OC_Q_KNN) {
      const double *qp = &j->queries[q * D];
      size_t want = (size_t)j->k + 8;
      for (;;) {
        if (tcap < want) {
          tcap = want;
          tmp = (uint32_t *)realloc(tmp, tcap * 4u);
        }
        size_t got = oc_knn_one(j->t, j->geom, qp, j->k, j->max_within, j->tol, tmp, tcap, &pq, &pqcap, &eb, &ecap);
        if (got <= tcap) {
          for (size_t i = 0; i < got; ++i) vec_push(&j->out, tmp[i]);
          break;
        }
        want = got;
      }
    } else if (j->kind == OC_Q_RANGE) {
      oc_range_one(j->t, j->geom, &j->queries[q * 2 * D], j->overlap_mode, j->tol, &j->out, &stack, &scap);
    } else if (j->kind == OC_Q_PICK) {
      oc_pick_one(j->t, j->geom, &j->queries[q * D], j->tol, &j->out, &stack, &scap);
    } else if (j->kind == OC_Q_PLANES) {
      oc_planes_one(j->t, j->geom, &j->queries[q * (size_t)j->nplanes * (D + 1)], j->nplanes, j->overlap_mode, j->tol, &j->out,
                    &stack, &scap);
    } else if (j->kind == OC_Q_RAY_ALL) {
      const double *qr = &j->queries[q * 2 * D];
      oc_ray_all_one(j->t, j->geom, qr, qr + D, j->tol, j->inc, j->max_dist, &j->out, &stack, &scap);
    } else {
      const double *qr = &j->queries[q * 2 * D];
      oc_ray_one(j->t, j->geom, qr, qr + D, j->tol, j->inc, j->max_dist, &j->out, &pq, &pqcap);
    }
    j->counts[q] = j->out.n - before;
  }
  free(pq);
  free(eb);
  free(stack);
  free(tmp);
  return NULL;
}

static oc_result *oc_run_jobs(oc_job *proto, uint64_t nq, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if ((uint64_t)nthreads > nq) nthreads = nq ? (int)nq : 1;
  oc_job *jobs = (oc_job *)calloc((size_t)nthreads, sizeof(oc_job));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  uint64_t *counts = (uint64_t *)calloc(nq + 1, 8);
  for (int i = 0; i < nthreads; ++i) {
    jobs[i] = *proto;
    jobs[i].q0 = nq * (uint64_t)i / (uint64_t)nthreads;
    jobs[i].q1 = nq * (uint64_t)(i + 1) / (uint64_t)nthreads;
    jobs[i].counts = counts;
    if (nthreads > 1) pthread_create(&th[i], NULL, oc_job_run, &jobs[i]);
    else oc_job_run(&jobs[i]);
  }
  if (nthreads > 1)
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  oc_result *r = (oc_result *)calloc(1, sizeof(oc_result));
  r->nq = nq;
  r->offsets = (uint64_t *)malloc(8u * (nq + 1));
  r->offsets[0] = 0;
  for (uint64_t q = 0; q < nq; ++q) r->offsets[q + 1] = r->offsets[q] + counts[q];
  r->nids = r->offsets[nq];
  r->ids = (uint32_t *)malloc(4u * (r->nids ? r->nids : 1));
  size_t pos = 0;
  for (int i = 0; i < nthreads; ++i) {
    if (jobs[i].out.n) memcpy(&r->ids[pos], jobs[i].out.ids, 4u * jobs[i].out.n);
    pos += jobs[i].out.n;
    free(jobs[i].out.ids);
  }
  free(jobs);
  free(th);
  free(counts);
  return r;
}

/* queries: nq x D points.  max_within = DBL_MAX for no limit. */
OC_API oc_result *oc_knn(const oc_tree *t, const double *pts, uint64_t nq, const double *queries, uint32_t k, double max_within,
                         double tol, int nthreads) {
  oc_job p = {0};
  p.t = t;
  p.geom = pts;
  p.queries = queries;
  p.kind = OC_Q_KNN;
  p.k = k;
  p.max_within = max_within;
  p.tol = tol;
  return oc_run_jobs(&p, nq, nthreads);
}

/* ranges: nq x (min[D], max[D]) */
OC_API oc_result *oc_range(const oc_tree *t, const double *geom, uint64_t nq, const double *ranges, int overlap_mode, double tol,
                           int nthreads) {
  oc_job p = {0};
  p.t = t;
  p.geom = geom;
  p.queries = ranges;
  p.kind = OC_Q_RANGE;
  p.overlap_mode = overlap_mode;
  p.tol = tol;
  return oc_run_jobs(&p, nq, nthreads);
}

/* rays: nq x (origin[D], dir[D]) */
OC_API oc_result *oc_ray_first(const oc_tree *t, const double *geom, uint64_t nq, const double *rays, double tol, double inc,
                               double max_dist, int nthreads) {
  oc_job p = {0};
  p.t = t;
  p.geom = geom;
  p.queries = rays;
  p.kind = OC_Q_RAY;
  p.tol = tol;
  p.inc = inc;
  p.max_dist = max_dist;
  return oc_run_jobs(&p, nq, nthreads);
}

/* picks: nq x D points */
OC_API oc_result *oc_pick(const oc_tree *t, const double *geom, uint64_t nq, const double *picks, double tol, int nthreads) {
  oc_job p = {0};
  p.t = t;
  p.geom = geom;
  p.queries = picks;
  p.kind = OC_Q_PICK;
  p.tol = tol;
  return oc_run_jobs(&p, nq, nthreads);
}

/* planes: nq x nplanes x (origo distance, normal[D]); mode 0 PlaneSearch, 1
 * PlanePositiveSegmentation (nplanes 1), 2 FrustumCulling */
OC_API oc_result *oc_planes(const oc_tree *t, const double *geom, uint64_t nq, uint32_t nplanes, const double *planes, int mode,
                            double tol, int nthreads) {
  oc_job p = {0};
  p.t = t;
  p.geom = geom;
  p.queries = planes;
  p.kind = OC_Q_PLANES;
  p.nplanes = nplanes;
  p.overlap_mode = mode;
  p.tol = tol;
  return oc_run_jobs(&p, nq, nthreads);
}

/* rays: nq x (origin[D], dir[D]); max_dist 0 = no limit (as the reference) */
OC_API oc_result *oc_ray_all(const oc_tree *t, const double *geom, uint64_t nq, const double *rays, double tol, double inc,
                             double max_dist, int nthreads) {
  oc_job p = {0};
  p.t = t;
  p.geom = geom;
  p.queries = rays;
  p.kind = OC_Q_RAY_ALL;
  p.tol = tol;
  p.inc = inc;
  p.max_dist = max_dist;
  return oc_run_jobs(&p, nq, nthreads);
}

/* ------------------------------------------------------------------------- */
/* Seeded generators shared by tests and the CPU baseline (exact doubles).     */
/* ------------------------------------------------------------------------- */
/* SplitMix64 (Steele, Lea, Flood 2014) */
static inline uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
OC_API void oc_splitmix_uniform(uint64_t seed, uint64_t n, double lo, double hi, double *out) {
  uint64_t s = seed;
  for (uint64_t i = 0; i < n; ++i) {
    double u = (double)(splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
    out[i] = lo + (hi - lo) * u;
  }
}
