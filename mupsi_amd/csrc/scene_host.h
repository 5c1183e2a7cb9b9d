// scene_host.h — host lowering of a mupsi_scene_desc into device descriptors.
//
// Every per-scene constant the kernels use (camera basis, tan of the field of
// view, GP normalisation, transform inverses, face normals, BVH) is computed
// here once on the host, in the same float operation order as the reference,
// so the device never re-derives them with a different math library.
#pragma once
#include <string>
#include <vector>

#include "../../include/mupsi.h"
#include "device.h"
#include "host_bvh.h"

namespace mupsi {

struct HostMesh {
  BvhResult bvh;
  std::vector<uint8_t> order;      // per node: axis | (lc<rc)<<2 | (lc>rc)<<3
  std::vector<float4> leaf_tris;   // 3 per leaf slot
  std::vector<float> vnormals;     // n_normals x 3 (may be empty)
  std::vector<int32_t> faces;      // n_faces x 3
  std::vector<float> fnormals;     // n_faces x 3
  std::vector<float> texcoords;    // n_texcoords x 2 (may be empty)
  std::vector<int32_t> face_tex;   // n_faces x 3 (may be empty)
  std::vector<float> verts;        // n_vertices x 3, local space
  std::vector<float> face_cdf;     // n_faces, cumulative area / total
  float total_area = 0.0f;
  dev::DevXf xf;
};

struct HostScene {
  dev::DevScene ds{};              // constants; device pointers are filled at upload
  std::vector<HostMesh> meshes;    // primitive meshes, indexed by DevPrim::mesh
  bool has_mean_mesh = false;
  HostMesh mean_mesh;
  std::vector<dev::DevPrim> prims;
  std::vector<std::vector<float>> prim_albedo;  // per primitive (empty = constant)
  std::vector<float> sky;
  std::vector<float> grid_mean, grid_var;  // ExternalGPMedium grids (res^3 each)
};

// Validate and lower; returns MUPSI_OK or an error status with *err set.
mupsi_status lower_scene(const mupsi_scene_desc& d, HostScene* out, std::string* err);
// A new GPMedium for a scene that has one (Controller::applyGpMedium,
// controller.cpp:397-426): updates the GP constants and mean of io->ds and the
// mean mesh; on failure io is left as it was.
mupsi_status lower_gp_medium(int32_t mean_type, const float* mean_center, float mean_radius,
                             const mupsi_mesh_desc* mean_mesh, const mupsi_gp_params& kernel,
                             HostScene* io, std::string* err);
bool validate_gp(const mupsi_gp_params& p, std::string* err);
dev::DevGP make_dev_gp(const mupsi_gp_params& p);
uint32_t host_constant_sampler_i(uint32_t seed);

// Message of the last failed call on this thread (mupsi_last_error).
extern thread_local std::string g_last_error;

}  // namespace mupsi
