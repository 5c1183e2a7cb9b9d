// host_bvh.cpp — BVH construction, reference topology (src/bvh/bvh.cpp:115-193).
//
// Kept alike on purpose: the split rule (longest axis of the node bounds,
// midpoint of the triangle-centroid range, std::partition, half split when
// degenerate) and the post-order node numbering define the traversal order
// and therefore which triangle id wins a closest-hit tie.  std::partition is
// the same libstdc++ algorithm the reference links, so the permutation of
// triangle indices inside each node is identical too.
#include "host_bvh.h"

#include <algorithm>
#include <cmath>

namespace mupsi {
namespace {

struct Builder {
  const float* v;
  const int32_t* f;
  BvhResult* out;

  float coord(int32_t vi, int axis) const { return v[3 * vi + axis]; }
  // Triangle centroid along one axis: (a + b + c) / 3 in float, left to right.
  float centroid(uint32_t ti, int axis) const {
    return (coord(f[3 * ti], axis) + coord(f[3 * ti + 1], axis) + coord(f[3 * ti + 2], axis)) /
           3.0f;
  }

  uint32_t build(uint32_t begin, uint32_t end) {
    std::vector<uint32_t>& idx = out->tri_index;
    const uint32_t count = end - begin;
    BvhNode node;
    for (int a = 0; a < 3; ++a) {
      node.bmin[a] = INFINITY;
      node.bmax[a] = -INFINITY;
    }
    for (uint32_t i = begin; i < end; ++i) {
      const uint32_t ti = idx[i];
      for (int k = 0; k < 3; ++k) {
        const int32_t vi = f[3 * ti + k];
        for (int a = 0; a < 3; ++a) {
          node.bmin[a] = std::min(node.bmin[a], coord(vi, a));
          node.bmax[a] = std::max(node.bmax[a], coord(vi, a));
        }
      }
    }
    if (count <= 4) {
      node.left = begin;
      node.right = count | kLeafFlag;
      out->nodes.push_back(node);
      return static_cast<uint32_t>(out->nodes.size() - 1);
    }
    const float dx = node.bmax[0] - node.bmin[0];
    const float dy = node.bmax[1] - node.bmin[1];
    const float dz = node.bmax[2] - node.bmin[2];
    const int axis = (dx > dy && dx > dz) ? 0 : (dy > dz ? 1 : 2);

    float lo = INFINITY, hi = -INFINITY;
    for (uint32_t i = begin; i < end; ++i) {
      const float c = centroid(idx[i], axis);
      lo = std::min(lo, c);
      hi = std::max(hi, c);
    }
    const float split = (lo + hi) * 0.5f;
    auto mid = std::partition(idx.begin() + begin, idx.begin() + end,
                              [&](uint32_t ti) { return centroid(ti, axis) < split; });
    uint32_t m = static_cast<uint32_t>(mid - idx.begin());
    if (m == begin || m == end) m = begin + count / 2;

    const uint32_t l = build(begin, m);
    const uint32_t r = build(m, end);
    node.left = l;
    node.right = r;
    out->nodes.push_back(node);
    return static_cast<uint32_t>(out->nodes.size() - 1);
  }
};

}  // namespace

BvhResult build_bvh(const float* vertices, const int32_t* faces, uint32_t n_faces) {
  BvhResult res;
  if (n_faces == 0) return res;
  res.tri_index.resize(n_faces);
  for (uint32_t i = 0; i < n_faces; ++i) res.tri_index[i] = i;
  res.nodes.reserve(2 * static_cast<size_t>(n_faces));
  Builder b{vertices, faces, &res};
  res.root = b.build(0, n_faces);
  return res;
}

}  // namespace mupsi
