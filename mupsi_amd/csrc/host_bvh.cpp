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

namespace {
uint32_t leaf_ref(const BvhNode& n) {
  return kLeafFlag | (n.right & ~kLeafFlag) << kLeafCountShift | n.left;
}
}  // namespace

std::vector<uint8_t> bvh_order(const BvhResult& b) {
  const std::vector<BvhNode>& nodes = b.nodes;
  std::vector<uint8_t> order(nodes.size(), 0);
  for (size_t i = 0; i < nodes.size(); ++i) {
    const BvhNode& n = nodes[i];
    if ((n.right & kLeafFlag) || n.left >= nodes.size() || n.right >= nodes.size()) continue;
    const float dx = n.bmax[0] - n.bmin[0], dy = n.bmax[1] - n.bmin[1], dz = n.bmax[2] - n.bmin[2];
    const int axis = (dx > dy && dx > dz) ? 0 : (dy > dz ? 1 : 2);
    const BvhNode& l = nodes[n.left];
    const BvhNode& r = nodes[n.right];
    const float lc = (l.bmin[axis] + l.bmax[axis]) * 0.5f;
    const float rc = (r.bmin[axis] + r.bmax[axis]) * 0.5f;
    order[i] = (uint8_t)(axis | ((lc < rc) ? 4 : 0) | ((lc > rc) ? 8 : 0));
  }
  return order;
}

bool build_records(const BvhResult& b, const std::vector<uint8_t>& order, BvhRecords& out) {
  out = BvhRecords{};
  const size_t n = b.nodes.size();
  if (n == 0) return true;
  if (b.root >= n || order.size() != n || n > (size_t)INT32_MAX) return false;
  for (const BvhNode& nd : b.nodes)
    if ((nd.right & kLeafFlag) &&
        ((nd.right & ~kLeafFlag) > 7u || nd.left > kLeafSlotMask ||
         (nd.right & ~kLeafFlag) + (uint64_t)nd.left > (uint64_t)kLeafSlotMask + 1))
      return false;
  // node 0 is the first node the reference's post-order build creates: a leaf
  if (!(b.nodes[0].right & kLeafFlag)) return false;
  out.occ_ref = leaf_ref(b.nodes[0]);
  for (int a = 0; a < 3; ++a) {
    out.occ_min[a] = b.nodes[0].bmin[a];
    out.occ_max[a] = b.nodes[0].bmax[a];
  }
  const BvhNode& root = b.nodes[b.root];
  if (root.right & kLeafFlag) {
    out.root_ref = leaf_ref(root);
    return true;
  }
  // breadth-first numbering of the internal nodes
  std::vector<int32_t> rec(n, -1);
  std::vector<uint32_t> bfs{b.root};
  std::vector<int> level{1};
  rec[b.root] = 0;
  for (size_t q = 0; q < bfs.size(); ++q) {
    const BvhNode& nd = b.nodes[bfs[q]];
    for (uint32_t c : {nd.left, nd.right}) {
      if (c >= n || rec[c] != -1 || c == b.root) return false;
      if (b.nodes[c].right & kLeafFlag) {
        rec[c] = -2;  // reached once
        continue;
      }
      rec[c] = (int32_t)bfs.size();
      bfs.push_back(c);
      level.push_back(level[q] + 1);
    }
  }
  out.recs.resize(bfs.size());
  for (size_t q = 0; q < bfs.size(); ++q) {
    const BvhNode& nd = b.nodes[bfs[q]];
    const BvhNode& l = b.nodes[nd.left];
    const BvhNode& r = b.nodes[nd.right];
    BvhRec& R = out.recs[q];
    for (int a = 0; a < 3; ++a) {
      R.lmin[a] = l.bmin[a];
      R.lmax[a] = l.bmax[a];
      R.rmin[a] = r.bmin[a];
      R.rmax[a] = r.bmax[a];
    }
    R.l = (l.right & kLeafFlag) ? leaf_ref(l) : (uint32_t)rec[nd.left];
    R.r = (r.right & kLeafFlag) ? leaf_ref(r) : (uint32_t)rec[nd.right];
    R.order = order[bfs[q]];
    out.depth = std::max(out.depth, level[q]);
  }
  // next: a left child continues with its parent's right child, a right child
  // with whatever its parent continues with (breadth-first order visits parents first)
  out.recs[0].next = -1;
  for (size_t q = 0; q < bfs.size(); ++q) {
    const BvhRec& R = out.recs[q];
    if (!(R.l & kLeafFlag)) out.recs[R.l].next = (int32_t)q;
    if (!(R.r & kLeafFlag)) out.recs[R.r].next = R.next;
  }
  return true;
}


bool build_xbins(const BvhRecords& L, const float lbmin[3], const float lbmax[3], XBins& out,
                 size_t max_entries) {
  out = XBins{};
  if (L.recs.empty()) return false;
  struct Leaf {
    float bmin[3], bmax[3];
    uint32_t ref;
  };
  std::vector<Leaf> leaves;
  leaves.reserve(L.recs.size() + 1);
  for (const BvhRec& R : L.recs) {
    if (R.l & kLeafFlag) leaves.push_back({{R.lmin[0], R.lmin[1], R.lmin[2]}, {R.lmax[0], R.lmax[1], R.lmax[2]}, R.l});
    if (R.r & kLeafFlag) leaves.push_back({{R.rmin[0], R.rmin[1], R.rmin[2]}, {R.rmax[0], R.rmax[1], R.rmax[2]}, R.r});
  }
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(lbmin[a]) || !std::isfinite(lbmax[a]) || !(lbmin[a] <= lbmax[a])) return false;
  std::vector<float> ext;
  ext.reserve(leaves.size());
  for (const Leaf& f : leaves) {
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(f.bmin[a]) || !std::isfinite(f.bmax[a]) || !(f.bmin[a] <= f.bmax[a]))
        return false;
    ext.push_back(std::max(f.bmax[1] - f.bmin[1], f.bmax[2] - f.bmin[2]));
  }
  std::nth_element(ext.begin(), ext.begin() + ext.size() / 2, ext.end());
  const float med = ext[ext.size() / 2];
  const float span = std::max(lbmax[1] - lbmin[1], lbmax[2] - lbmin[2]);
  // bin coordinates and counts must stay far inside int64 (the walk serves the rest)
  if (!std::isfinite(span) || span > 1e15f) return false;
  const float base = std::max(med, span * 1e-4f);
  int e = base > 0.0f ? (int)std::floor(std::log2(base)) : 0;
  e = std::max(-60, std::min(60, e));
  const float y0 = lbmin[1], z0 = lbmin[2];
  for (; e <= 64; ++e) {
    const float inv = std::ldexp(1.0f, -e);
    // the device's bin coordinate: (y - y0) * inv, one rounding (inv a power of two)
    const auto bin = [&](float v, float v0) { return (int64_t)std::floor((v - v0) * inv); };
    const int64_t ny = bin(lbmax[1], y0) + 1, nz = bin(lbmax[2], z0) + 1;
    if (ny < 1 || nz < 1 || (double)ny * (double)nz > (double)max_entries) continue;
    size_t total = 0;
    for (const Leaf& f : leaves)
      total += (size_t)(bin(f.bmax[1], y0) - bin(f.bmin[1], y0) + 1) *
               (size_t)(bin(f.bmax[2], z0) - bin(f.bmin[2], z0) + 1);
    if (total > max_entries) continue;
    std::vector<uint32_t> count((size_t)(ny * nz) + 1, 0);
    for (const Leaf& f : leaves)
      for (int64_t iz = bin(f.bmin[2], z0); iz <= bin(f.bmax[2], z0); ++iz)
        for (int64_t iy = bin(f.bmin[1], y0); iy <= bin(f.bmax[1], y0); ++iy)
          ++count[(size_t)(iz * ny + iy) + 1];
    for (size_t b = 1; b < count.size(); ++b) count[b] += count[b - 1];
    out.offsets = count;
    out.entries.assign(total, XBinEntry{});
    std::vector<uint32_t> fill(count.begin(), count.end() - 1);
    for (const Leaf& f : leaves) {
      const XBinEntry E{f.bmax[0], f.bmin[1], f.bmax[1], f.bmin[2], f.bmax[2], f.bmin[0], f.ref, 0u};
      for (int64_t iz = bin(f.bmin[2], z0); iz <= bin(f.bmax[2], z0); ++iz)
        for (int64_t iy = bin(f.bmin[1], y0); iy <= bin(f.bmax[1], y0); ++iy)
          out.entries[fill[(size_t)(iz * ny + iy)]++] = E;
    }
    for (size_t b = 0; b + 1 < out.offsets.size(); ++b)
      std::sort(out.entries.begin() + out.offsets[b], out.entries.begin() + out.offsets[b + 1],
                [](const XBinEntry& a, const XBinEntry& c) {
                  return a.xmax > c.xmax || (a.xmax == c.xmax && a.ref < c.ref);
                });
    out.ny = (int)ny;
    out.nz = (int)nz;
    out.y0 = y0;
    out.z0 = z0;
    out.inv = inv;
    out.log2_side = e;
    return true;
  }
  out = XBins{};
  return false;
}

}  // namespace mupsi
