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

// Per node, the near/far order BVH::intersect pushes an internal node's children
// in (bvh.cpp:228-245): axis | (lc < rc) << 2 | (lc > rc) << 3, with axis the
// longest axis of the node's bounds and lc, rc the children's box centres on
// it; 0 for leaves.
std::vector<uint8_t> bvh_order(const BvhResult& b);

// The device layout of the same tree: one 64-byte record per internal node
// holding the bounds of both its children side by side, so that a traversal
// step reads the two boxes it tests with one record load (a walk that tests one
// child at a time reads the 32-byte half of that child).  Records are
// numbered breadth-first from the root (record 0), so the top levels are the
// first records (the part a kernel may copy into LDS).  A child reference is a
// record index, or for a leaf kLeafFlag | count << 28 | first triangle slot.
struct BvhRec {
  // first half: the left child
  float lmin[3], lmax[3];  // its bounds
  uint32_t l;              // its reference
  uint32_t order;          // axis | (lc < rc) << 2 | (lc > rc) << 3 (near/far order)
  // second half: the right child
  float rmin[3], rmax[3];
  uint32_t r;
  int32_t next;            // the record whose right child a left-first pre-order
                           // walk visits once this subtree is done (-1: none)
};
static_assert(sizeof(BvhRec) == 64, "BvhRec must stay 64 bytes");

constexpr uint32_t kLeafCountShift = 28;
constexpr uint32_t kLeafSlotMask = (1u << kLeafCountShift) - 1u;

struct BvhRecords {
  std::vector<BvhRec> recs;
  uint32_t root_ref = 0;   // reference of the root (a leaf for meshes of <= 4 triangles)
  uint32_t occ_ref = 0;    // reference of node 0, where BVH::occluded starts (a leaf)
  float occ_min[3] = {0, 0, 0}, occ_max[3] = {0, 0, 0};
  int depth = 0;           // levels of internal nodes on the longest root-to-leaf path
};

// Records of b (order: the near/far byte of each internal node, indexed like
// b.nodes).  False if b is not a tree rooted at b.root with every node reached
// once, or a leaf slot does not fit the reference encoding.
bool build_records(const BvhResult& b, const std::vector<uint8_t>& order, BvhRecords& out);

// Leaves binned by their y-z bounds, for Mesh::inside's +x parity ray: the
// local-space y-z box of the mesh is cut into square bins of side 2^e (a power
// of two, so that a bin coordinate is (y - y0) * 2^-e with one rounding, in the
// subtraction, and is monotone in y); every leaf is listed in every bin its y-z
// bounds reach.  A point whose y, z lie inside a leaf's bounds then finds that
// leaf in its own bin, the host and the device computing bin coordinates with
// the same float operations.  A bin's entries are sorted by the leaf's max x,
// largest first, so that a +x ray from x stops at the first leaf ending before x.
struct XBinEntry {  // 32 bytes, read as two float4
  float xmax, ymin, ymax, zmin;
  float zmax, xmin;
  uint32_t ref;  // the leaf's reference (kLeafFlag | count << 28 | first slot)
  uint32_t pad;
};
static_assert(sizeof(XBinEntry) == 32, "XBinEntry must stay 32 bytes");

struct XBins {
  std::vector<XBinEntry> entries;
  std::vector<uint32_t> offsets;  // ny * nz + 1: bin (iy, iz) holds [off[b], off[b + 1]), b = iz * ny + iy
  int ny = 0, nz = 0;
  float y0 = 0.f, z0 = 0.f, inv = 0.f;  // inv = 2^-e
  int log2_side = 0;
};

// Bins of the leaves of L (a tree of internal nodes; false for a single leaf,
// non-finite bounds, or more than max_entries entries at every side tried).
// The side starts at the power of two at or below the median leaf extent in y
// and z and doubles while the bins or entries exceed max_entries.
bool build_xbins(const BvhRecords& L, const float lbmin[3], const float lbmax[3], XBins& out,
                 size_t max_entries = 8u << 20);

}  // namespace mupsi
