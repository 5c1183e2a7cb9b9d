// host_bvh.h — host-side BVH construction with the reference topology.
//
// The device traverses exactly the node array mupsi's BVH::build produces
// (post-order, root last, leaves of <= 4 triangles split at the centroid
// midpoint of the longest axis; src/bvh/bvh.cpp:115-193), so that closest-hit
// ties resolve to the same triangle id as the CPU renderer.
#pragma once
#include <cstdint>
#include <vector>

namespace mupsi {

struct BvhNode {  // 32 bytes, the layout of BVH::Node (bvh.h:26-36)
  float bmin[3];
  float bmax[3];
  uint32_t left;   // leaf: first slot in tri_index; internal: left child
  uint32_t right;  // leaf: count | 0x80000000; internal: right child
};
static_assert(sizeof(BvhNode) == 32, "BvhNode must stay 32 bytes");

constexpr uint32_t kLeafFlag = 0x80000000u;

struct BvhResult {
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> tri_index;  // leaf slot -> triangle id
  uint32_t root = 0;
};

// vertices: n_vertices x 3 floats; faces: n_faces x 3 vertex indices.
BvhResult build_bvh(const float* vertices, const int32_t* faces, uint32_t n_faces);

}  // namespace mupsi
