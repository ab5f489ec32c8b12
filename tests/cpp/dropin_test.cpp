// Exercises include/otg/orthotree_gpu.hpp the way reference user code calls the
// static trees (core/ot_aliases.h names, PAR_EXEC constructor, query methods),
// on the reference's own known-answer data (tests/unittests/general.tests.cpp).
// Prints "DROPIN OK" and exits 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "otg/orthotree_gpu.hpp"

using namespace otg;

static int failures = 0;
#define STEP(msg)                   \
  do {                              \
    std::printf("step: %s\n", msg); \
    std::fflush(stdout);            \
  } while (0)
#define EXPECT(cond)                                                  \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static bool same_set(std::vector<index_t> a, std::vector<index_t> b) {
  std::sort(a.begin(), a.end());
  std::sort(b.begin(), b.end());
  return a == b;
}

int main() {
  // getPointSetNo1<2>() (general.tests.cpp:72-130), tree (points, 3, nullopt, 3)
  std::vector<Point2D> pts = {{0.0, 0.0}, {3.0, 3.0}, {5.0, 3.0}, {5.2, 3.0}, {7.0, 1.0}, {0.0, 5.0}, {1.0, 5.0},
                              {0.0, 7.0}, {1.0, 7.0}, {3.0, 7.0}, {3.0, 7.0}, {2.2, 4.5}, {2.5, 4.5}, {3.4, 4.5},
                              {3.6, 4.6}, {2.4, 5.5}, {2.5, 5.5}, {3.5, 5.5}, {3.5, 5.5}, {5.0, 5.0}, {5.0, 5.0},
                              {7.0, 5.0}, {7.0, 5.0}, {6.5, 6.5}, {6.5, 6.5}, {7.5, 6.5}, {7.5, 6.5}, {6.5, 7.5},
                              {6.5, 7.5}, {7.5, 7.5}, {7.5, 7.5}, {8.0, 8.0}};
  STEP("build quadtree points");
  StaticQuadtreePoint qt(PAR_EXEC, pts, 3, std::nullopt, 3);
  STEP("node count");
  EXPECT(qt.GetNodeCount() > 1);
  EXPECT(qt.GetMaxDepthID() == 3);
  STEP("knn");
  EXPECT(same_set(qt.GetNearestNeighbors({3.5, 5.5}, 3, pts), {17, 18, 14}));            // :2488
  EXPECT(same_set(qt.GetNearestNeighbors({3.5, 5.5}, 4, 0.0, pts), {17, 18}));           // :2510
  EXPECT(same_set(qt.GetNearestNeighbors({-1.0, -1.0}, 5, pts), {0, 1, 5, 6, 11}));      // :2593
  EXPECT(same_set(qt.GetNearestNeighbors({2.0, 2.0}, 2, pts), {1, 11}));                 // :2558
  {  // :2532 -- k = 100 over 32 points: every point, nearest first
    auto all = qt.GetNearestNeighbors({3.5, 5.5}, 100, pts);
    std::vector<index_t> every(pts.size());
    for (size_t i = 0; i < every.size(); ++i) every[i] = (index_t)i;
    EXPECT(all.size() == pts.size() && same_set(all, every));
  }
  STEP("node api");
  // node API: children ranges cover every node once, entities cover every id once
  {
    std::vector<int> seen(qt.GetNodeCount(), 0), ent(pts.size(), 0);
    seen[0] = 1;
    for (uint32_t i = 0; i < qt.GetNodeCount(); ++i) {
      auto [b, e] = qt.GetNodeChildren(i);
      for (auto c = b; c < e; ++c) ++seen[c];
      for (auto id : qt.GetNodeEntities(i)) ++ent[id];
    }
    EXPECT(std::all_of(seen.begin(), seen.end(), [](int x) { return x == 1; }));
    EXPECT(std::all_of(ent.begin(), ent.end(), [](int x) { return x == 1; }));
  }
  STEP("range");
  auto r = qt.RangeSearch(BoundingBox2D{{2.0, 4.0}, {4.0, 6.0}}, pts);
  std::vector<index_t> expect_r;
  for (index_t i = 0; i < pts.size(); ++i)
    if (pts[i][0] > 2.0 - BASE_TOLERANCE && pts[i][0] < 4.0 + BASE_TOLERANCE && pts[i][1] > 4.0 - BASE_TOLERANCE &&
        pts[i][1] < 6.0 + BASE_TOLERANCE)
      expect_r.push_back(i);
  EXPECT(same_set(r, expect_r));

  // ray KATs on the 4-box diagonal (general.tests.cpp:1360-1453), QuadtreeBox(boxes, 3, nullopt, 2)
  std::vector<BoundingBox2D> boxes = {{{0, 0}, {1, 1}}, {{1, 1}, {2, 2}}, {{2, 2}, {3, 3}}, {{3, 3}, {4, 4}}};
  STEP("box tree + rays");
  StaticQuadtreeBox qb(PAR_EXEC, boxes, 3, std::nullopt, 2);
  EXPECT((qb.RayIntersectedFirst({0.0, 0.5}, {1.0, 0.0}, boxes, 0.0) == std::vector<index_t>{0}));
  EXPECT((qb.RayIntersectedFirst({0.0, 2.0}, {1.0, 0.0}, boxes, 0.0) == std::vector<index_t>{2, 1}));
  EXPECT((qb.RayIntersectedFirst({1.5, 1.0}, {-1.0, 0.0}, boxes, 0.0) == std::vector<index_t>{0, 1}));
  EXPECT(qb.RayIntersectedFirst({0.0, -1.0}, {1.0, 0.0}, boxes, 0.0).empty());
  EXPECT((qb.RayIntersectedFirst(RayND<2>{{3.5, 3.5}, {-1.0 / std::sqrt(5.0), -2.0 / std::sqrt(5.0)}}, boxes, 0.0) ==
          std::vector<index_t>{3}));

  // collision: overlapping 3D boxes (strict, with tolerance), compared with brute force
  std::vector<BoundingBox3D> bx;
  for (int i = 0; i < 200; ++i) {
    double x = (i * 37 % 100) * 0.1, y = (i * 53 % 100) * 0.1, z = (i * 71 % 100) * 0.1;
    bx.push_back({{x, y, z}, {x + 0.8, y + 0.6, z + 0.7}});
  }
  STEP("collision");
  StaticOctreeBox ob(PAR_EXEC, bx, 5, std::nullopt, 4);
  STEP("collision: built");
  std::printf("nodes %zu\n", ob.GetNodeCount());
  auto pairs = ob.CollisionDetection<ParExec>(bx);
  STEP("collision: detected");
  std::vector<std::pair<index_t, index_t>> brute;
  for (index_t i = 0; i < bx.size(); ++i)
    for (index_t j = i + 1; j < bx.size(); ++j) {
      bool ov = true;
      for (int d = 0; d < 3; ++d)
        ov = ov && (bx[i].Min[d] + BASE_TOLERANCE < bx[j].Max[d]) && (bx[j].Min[d] + BASE_TOLERANCE < bx[i].Max[d]);
      if (ov) brute.push_back({i, j});
    }
  EXPECT(pairs == brute);

  // batch overload
  std::vector<Point3D> p3;
  for (int64_t i = 0; i < 5000; ++i)
    p3.push_back({double(i * 7919 % 1000) * 0.1, double(i * 104729 % 1000) * 0.1, double(i * 1299709 % 1000) * 0.1});
  STEP("batch knn");
  StaticOctreePoint op(PAR_EXEC, p3, 6, BoundingBox3D{{0, 0, 0}, {100, 100, 100}}, 16);
  std::vector<Point3D> q = {{50, 50, 50}, {1, 2, 3}, {99, 0, 42}};
  auto br = op.GetNearestNeighbors(std::span<const Point3D>(q), 8);
  EXPECT(br.size() == 3);
  for (size_t i = 0; i < q.size(); ++i) {
    auto one = op.GetNearestNeighbors(q[i], 8, p3);
    EXPECT(std::vector<index_t>(br[i].begin(), br[i].end()) == one);
  }
  // box-tree kNN with MinMaxDistance (general.tests.cpp:2648, :2772, :2788)
  STEP("box knn");
  std::vector<BoundingBox2D> mm = {{{+0.20, 0.00}, {+2.31, 2.31}}, {{+1.00, 0.00}, {+1.01, 1.41}},
                                   {{+0.40, 0.00}, {+0.89, 0.89}}, {{+2.10, 0.00}, {+3.66, 3.66}},
                                   {{-2.21, 0.00}, {-0.20, 2.21}}, {{-2.50, 0.00}, {-0.90, 2.47}}};
  StaticQuadtreeBox qm(PAR_EXEC, mm, 3, std::nullopt, 2);
  EXPECT(same_set(qm.GetNearestNeighbors({0.0, 0.0}, 1, std::numeric_limits<double>::max(), mm,
                                         std::numeric_limits<double>::epsilon()),
                  {0, 2, 4}));
  std::vector<BoundingBox2D> b3 = {{{0.0, 0.0}, {1.0, 1.0}}, {{1.0, 1.0}, {3.0, 3.0}}, {{4.0, 4.0}, {5.0, 5.0}}};
  StaticQuadtreeBox q3(PAR_EXEC, b3, 3, std::nullopt, 2);
  EXPECT(same_set(q3.GetNearestNeighbors({2.0, 2.0}, 1, b3), {1}));
  EXPECT(same_set(q3.GetNearestNeighbors({-1.0, 1.0}, 1, b3), {0}));
  // PickSearch KATs (general.tests.cpp:2928, :2947): QuadtreeBox(boxes, 3, nullopt, 2)
  STEP("pick search");
  std::vector<BoundingBox2D> b5 = {{{0, 0}, {1, 1}}, {{1, 1}, {2, 2}}, {{2, 2}, {3, 3}}, {{3, 3}, {4, 4}}, {{1.2, 1.2}, {2.8, 2.8}}};
  StaticQuadtreeBox q5(PAR_EXEC, b5, 3, std::nullopt, 2);
  EXPECT(same_set(q5.PickSearch({2.5, 2.5}, b5), {2, 4}));
  EXPECT(same_set(q5.PickSearch({2.0, 2.0}, b5), {1, 2, 4}));
  // RayIntersectedAll KATs (general.tests.cpp:1545, :1639-1669)
  STEP("ray all");
  std::vector<BoundingBox2D> r5 = {{{2.0, 0.0}, {3.0, 1.0}}, {{3.0, 1.0}, {4.0, 2.0}}, {{3.0, 2.0}, {4.0, 3.0}},
                                   {{2.0, 3.0}, {4.0, 4.0}}, {{2.5, 2.5}, {3.5, 3.5}}};
  StaticQuadtreeBox qr(PAR_EXEC, r5, 3, std::nullopt, 2);
  const double h = 1.4142135623730951 * 0.5;
  EXPECT((qr.RayIntersectedAll({2.7, 2.6}, {h, h}, r5, 0.0) == std::vector<index_t>{4, 2, 3}));
  EXPECT(same_set(q5.RayIntersectedAll({1.1, 2.0}, {1.0, 0.0}, b5, 0.0), {1, 4, 2}));
  EXPECT(same_set(q5.RayIntersectedAll({1.0, 1.9}, {1.0, 0.0}, b5, 0.901), {0, 1, 4, 2}));
  // two-tree collision (general.tests.cpp:3081): { {3, 3}, {2, 4}, {3, 4} }
  STEP("two-tree collision");
  StaticQuadtreeBox ql(PAR_EXEC, boxes, 3, std::nullopt, 2);
  auto dual = StaticQuadtreeBox::CollisionDetection(ql, boxes, qr, r5);
  std::sort(dual.begin(), dual.end());
  EXPECT((dual == std::vector<std::pair<index_t, index_t>>{{2, 4}, {3, 3}, {3, 4}}));
  EXPECT((ql.CollisionDetection(boxes, qr, r5) == dual));
  // plane KATs: PlaneIntersection_3D_XY_d0 (general.tests.cpp:2977), FrustumCulling (adapter.tests.cpp:283)
  STEP("planes");
  std::vector<BoundingBox3D> pb = {{{0.0, 0.0, 0.0}, {1.0, 1.0, 0.0}}, {{1.0, 1.0, -1.0}, {2.0, 2.0, 0.0}},
                                   {{2.0, 2.0, -1.0}, {3.0, 3.0, -0.2}}, {{3.0, 3.0, 1.0}, {4.0, 4.0, 1.2}},
                                   {{1.2, 1.2, -1.0}, {2.8, 2.8, 1.0}}};
  StaticOctreeBox opb(PAR_EXEC, pb, 3, std::nullopt, 2);
  EXPECT(same_set(opb.PlaneIntersection(0.0, Point3D{0.0, 0.0, 1.0}, pb, 0.01), {0, 1, 4}));
  StaticQuadtreeBox qf(PAR_EXEC, b5, 3, std::nullopt, 1);
  const double s2 = std::sqrt(2.0);
  std::vector<PlaneND<2>> fr = {{2 * s2, {1.0 / s2, 1.0 / s2}}, {-2.0, {0.0, -1.0}}};
  EXPECT(same_set(qf.FrustumCulling(fr, b5, double(0.01f)), {1, 2, 4}));
  // managed (entity-owning) trees: PlaneSearchC_3D_XZ_dP4 (general.tests.cpp:2458, OctreePointM) and
  // PlaneIntersectionC_3D_YZ_dN1 (:3048, OctreeBoxM)
  STEP("managed trees");
  std::vector<Point3D> vpt = {{2.0, -1.0, 0.0}, {3.0, -2.0, 1.0}, {4.0, -3.0, 0.02}, {4.0, -2.0, 0.00}, {5.0, 0.0, 1.00999}};
  StaticOctreePointM mpt(vpt, 3, std::nullopt, 2);
  EXPECT(same_set(mpt.PlaneSearch(4.0, Point3D{1.0, 0.0, 0.0}, 0.01), {2, 3}));
  EXPECT(mpt.GetData().size() == vpt.size() && mpt.Get(4)[2] == 1.00999);
  EXPECT(same_set(mpt.GetNearestNeighbors({4.0, -2.5, 0.0}, 2), {2, 3}));
  std::vector<BoundingBox3D> mbx = {{{0.0, 0.0, 0.0}, {1.0, 1.0, 0.0}}, {{1.0, 1.0, -1.0}, {2.0, 2.0, 0.0}},
                                    {{2.0, 2.0, -1.0}, {3.0, 3.0, -0.2}}, {{3.0, 3.0, 1.0}, {4.0, 4.0, 1.2}},
                                    {{1.2, 1.2, -1.1}, {2.8, 2.8, 1.1}}};
  auto mbox = StaticOctreeBoxM::Create(mbx, 3, std::nullopt, 2);
  EXPECT(mbox.PlaneIntersection(-1.0, Point3D{1.0, 0.0, 0.0}, 0.01).empty());
  EXPECT(same_set(mbox.PlaneIntersection(-1.0, Point3D{0.0, 0.0, 1.0}, 0.01), {1, 2, 4}));
  EXPECT((mbox.CollisionDetection() == std::vector<std::pair<index_t, index_t>>{{1, 4}, {2, 4}}));
  StaticQuadtreeBoxM ml(PAR_EXEC, boxes, 3, std::nullopt, 2), mr(PAR_EXEC, r5, 3, std::nullopt, 2);
  auto mdual = StaticQuadtreeBoxM::CollisionDetection(ml, mr);
  std::sort(mdual.begin(), mdual.end());
  EXPECT((mdual == std::vector<std::pair<index_t, index_t>>{{2, 4}, {3, 3}, {3, 4}}));
  // archives (serialization.h:373-386, :433-439) of a tree small enough to
  // derive by hand: root + three one-point leaves at depth 2, NodeStorage256
  STEP("archives");
  std::vector<Point2D> hp = {{0.5, 0.5}, {3.5, 0.5}, {0.5, 3.5}};
  const BoundingBox2D hbox{{0.0, 0.0}, {4.0, 4.0}};
  StaticQuadtreePoint ht(PAR_EXEC, hp, 2, hbox, 1);
  std::vector<uint8_t> hand;
  auto put = [&hand](const auto &v) {
    const auto *b = reinterpret_cast<const uint8_t *>(&v);
    hand.insert(hand.end(), b, b + sizeof(v));
  };
  auto arr = [&](double x, double y) {
    put(uint32_t(2));
    put(x);
    put(y);
  };
  put(uint32_t(0));  // core and base SERIALIZED_VERSION_ID
  put(uint32_t(0));
  put(uint64_t(1));  // m_maxElementNum
  put(uint32_t(2));  // m_maxDepthID
  put(uint32_t(4));  // m_nodeSize
  arr(4, 4), arr(2, 2), arr(1, 1), arr(0.5, 0.5);
  put(uint32_t(0));  // variant index: NodeStorage256
  put(uint32_t(2)), put(uint8_t(1)), put(uint8_t(4));  // nodeChildSegmentBegins
  put(uint32_t(4));                                    // nodeEntitySegment
  for (uint8_t b : {0, 0, 0, 1, 1, 1, 2, 1}) put(b);
  put(uint32_t(4));  // m_nodeDepthIDs
  for (uint8_t b : {0, 2, 2, 2}) put(b);
  put(uint32_t(3));  // m_entityStorage
  for (uint32_t i : {0u, 1u, 2u}) put(i);
  put(uint32_t(4));  // m_nodeGeometry (min points)
  arr(0, 0), arr(0, 0), arr(3, 0), arr(0, 3);
  EXPECT(ht.Serialize() == hand);
  StaticQuadtreePointM hm(hp, 2, hbox, 1);
  put(uint32_t(3));  // the managed tree's entity vector
  for (const Point2D &p : hp) arr(p[0], p[1]);
  EXPECT(hm.Serialize() == hand);

  // loading archives back (serialization.h:373-386, :433-439): the hand archive into a
  // core and a managed tree, the reference's TestStaticOctreeBinary round trip
  // (tests/unittests/serialize.tests.cpp:133-146), and a rejected mismatch
  StaticQuadtreePointM lm;
  EXPECT(lm.Deserialize(hand));
  EXPECT(lm.GetData().size() == 3 && lm.Get(2)[1] == hp[2][1]);
  EXPECT(lm.Serialize() == hand);
  EXPECT(lm.GetNearestNeighbors(Point2D{3.4, 0.6}, 1).size() == 1 && lm.GetNearestNeighbors(Point2D{3.4, 0.6}, 1)[0] == 1);
  {
    const std::vector<Point3D> two = {{0.1, 0.1, 0.1}, {0.9, 0.9, 0.9}};
    StaticOctreePointM saved(two, 3, BoundingBox3D{{0, 0, 0}, {1, 1, 1}});
    StaticOctreePointM loaded;
    EXPECT(loaded.Deserialize(saved.Serialize()));
    auto a = saved.RangeSearch(BoundingBox3D{{0, 0, 0}, {1, 1, 1}});
    auto b = loaded.RangeSearch(BoundingBox3D{{0, 0, 0}, {1, 1, 1}});
    EXPECT(a.size() == 2 && a == b);
    StaticOctreePoint core;
    EXPECT(core.Deserialize(saved.GetCore().Serialize(), two));
    const std::vector<Point3D> moved = {{0.1, 0.1, 0.1}, {1.5, 0.9, 0.9}};  // outside the root
    EXPECT(!core.Deserialize(saved.GetCore().Serialize(), moved));
  }
  if (failures == 0) std::printf("DROPIN OK\n");
  return failures ? 1 : 0;
}
