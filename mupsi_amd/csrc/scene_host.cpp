// scene_host.cpp — host lowering of the scene description (see scene_host.h).
#include "scene_host.h"

#include <cmath>
#include <cstring>

namespace mupsi {

thread_local std::string g_last_error;

namespace {

struct V {
  float x, y, z;
};
V vsub(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
float vsq(V v) { return std::fmaf(v.x, v.x, std::fmaf(v.y, v.y, v.z * v.z)); }
V vcross(V a, V b) {
  return V{std::fmaf(a.y, b.z, -(a.z * b.y)), std::fmaf(a.z, b.x, -(a.x * b.z)),
           std::fmaf(a.x, b.y, -(a.y * b.x))};
}
V vnormalized(V v) {  // Eigen normalized()
  const float z = vsq(v);
  if (!(z > 0.0f)) return v;
  const float s = std::sqrt(z);
  return V{v.x / s, v.y / s, v.z / s};
}
V vld(const float* p) { return V{p[0], p[1], p[2]}; }

// Translation(pos) * Scaling(s); inverse fixed as 1/s, -pos/s (via -pos * (1/s)).
dev::DevXf make_xf(const float pos[3], float s) {
  dev::DevXf x{};
  x.s = s;
  x.inv_s = 1.0f / s;
  for (int a = 0; a < 3; ++a) {
    x.pos[a] = pos[a];
    x.inv_t[a] = -(pos[a] * x.inv_s);
  }
  // Eigen isIdentity() with precision 1e-5 on the inverse matrix.
  const float prec = 1e-5f;
  bool ident = std::fabs(x.inv_s - 1.0f) <= prec * std::fmin(std::fabs(x.inv_s), 1.0f);
  for (int a = 0; a < 3; ++a) ident = ident && std::fabs(x.inv_t[a]) <= prec;
  x.identity = ident ? 1 : 0;
  return x;
}

uint32_t hpcg(uint64_t& s) {
  const uint64_t old = s;
  s = old * 6364136223846793005ULL + 1ULL;
  const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((uint32_t)(-(int32_t)rot) & 31u));
}

bool lower_mesh(const mupsi_mesh_desc& d, HostMesh* m, std::string* err) {
  if (d.n_faces <= 0 || d.n_vertices <= 0 || !d.vertices || !d.faces) {
    *err = "mesh has no geometry";
    return false;
  }
  if (!(d.scale != 0.0f) || !std::isfinite(d.scale)) {
    *err = "mesh scale must be finite and non-zero";
    return false;
  }
  for (int64_t i = 0; i < 3LL * d.n_faces; ++i) {
    if (d.faces[i] < 0 || d.faces[i] >= d.n_vertices) {
      *err = "mesh face index out of range";
      return false;
    }
    if (d.face_tex && d.texcoords && (d.face_tex[i] < 0 || d.face_tex[i] >= d.n_texcoords)) {
      *err = "mesh texcoord index out of range";
      return false;
    }
  }
  if (d.normals && d.n_normals > 0 && d.n_normals < d.n_vertices) {
    // Mesh::interpolateNormal reads normals_[vertex index]; fewer vn lines than
    // v lines would read past the array in the reference.
    *err = "mesh has fewer normals than vertices";
    return false;
  }
  m->faces.assign(d.faces, d.faces + 3LL * d.n_faces);
  if (d.normals && d.n_normals > 0) m->vnormals.assign(d.normals, d.normals + 3LL * d.n_normals);
  if (d.texcoords && d.n_texcoords > 0 && d.face_tex) {
    m->texcoords.assign(d.texcoords, d.texcoords + 2LL * d.n_texcoords);
    m->face_tex.assign(d.face_tex, d.face_tex + 3LL * d.n_faces);
  }
  m->xf = make_xf(d.position, d.scale);
  // Mesh::computeFaceNormals (mesh.cpp:140-154)
  m->fnormals.resize(3LL * d.n_faces);
  for (int i = 0; i < d.n_faces; ++i) {
    const V v0 = vld(d.vertices + 3LL * d.faces[3 * i]);
    const V v1 = vld(d.vertices + 3LL * d.faces[3 * i + 1]);
    const V v2 = vld(d.vertices + 3LL * d.faces[3 * i + 2]);
    const V n = vcross(vsub(v1, v0), vsub(v2, v0));
    const float len = std::sqrt(vsq(n));
    const bool ok = len > 1e-12f;
    m->fnormals[3 * i] = ok ? n.x / len : 0.0f;
    m->fnormals[3 * i + 1] = ok ? n.y / len : 0.0f;
    m->fnormals[3 * i + 2] = ok ? n.z / len : 0.0f;
  }
  // Mesh::prepareForRender (mesh.cpp:298-319): face areas 0.5 |e1 x e2|, the
  // float running total, then every entry divided by the total.
  m->verts.assign(d.vertices, d.vertices + 3LL * d.n_vertices);
  m->face_cdf.resize(d.n_faces);
  float total = 0.0f;
  for (int i = 0; i < d.n_faces; ++i) {
    const V v0 = vld(d.vertices + 3LL * d.faces[3 * i]);
    const V v1 = vld(d.vertices + 3LL * d.faces[3 * i + 1]);
    const V v2 = vld(d.vertices + 3LL * d.faces[3 * i + 2]);
    total += 0.5f * std::sqrt(vsq(vcross(vsub(v1, v0), vsub(v2, v0))));
    m->face_cdf[i] = total;
  }
  if (total > 0.0f)
    for (float& c : m->face_cdf) c /= total;
  m->total_area = total;
  m->bvh = build_bvh(d.vertices, d.faces, (uint32_t)d.n_faces);
  m->order = bvh_order(m->bvh);
  m->leaf_tris.resize(3 * m->bvh.tri_index.size());
  for (size_t s = 0; s < m->bvh.tri_index.size(); ++s) {
    const uint32_t ti = m->bvh.tri_index[s];
    float idf;
    std::memcpy(&idf, &ti, 4);
    for (int k = 0; k < 3; ++k) {
      const float* v = d.vertices + 3LL * d.faces[3 * ti + k];
      m->leaf_tris[3 * s + k] = make_float4(v[0], v[1], v[2], k == 0 ? idf : 0.0f);
    }
  }
  return true;
}

bool texture_ok(const mupsi_texture_desc& t) { return t.rgb && t.width > 0 && t.height > 0; }

}  // namespace

uint32_t host_constant_sampler_i(uint32_t seed) {
  uint64_t s = seed == 0 ? 1ULL : seed;
  for (int i = 0; i < 4; ++i) hpcg(s);
  return hpcg(s);
}

bool validate_gp(const mupsi_gp_params& p, std::string* err) {
  if (!(p.kernel_radius > 0.0f) || !std::isfinite(p.kernel_radius)) {
    *err = "kernel.length_scale (cell size) must be positive";
    return false;
  }
  if (p.points_per_cell < 1 || p.points_per_cell > 64) {
    *err = "kernel.points_per_cell must be in 1..64";
    return false;
  }
  const float L2 = vsq(V{p.aniso[0], p.aniso[1], p.aniso[2]});
  if (!(L2 > 0.0f) || !std::isfinite(L2)) {
    *err = "kernel.aniso must be non-zero";
    return false;
  }
  if (!std::isfinite(p.sigma)) {
    *err = "kernel.sigma must be finite";
    return false;
  }
  return true;
}

dev::DevGP make_dev_gp(const mupsi_gp_params& p) {
  dev::DevGP g{};
  g.sigma = p.sigma;
  g.kr = p.kernel_radius;
  g.L2 = vsq(V{p.aniso[0], p.aniso[1], p.aniso[2]});
  g.ppc = p.points_per_cell;
  // SparseGPNoiseGenerator::Var() (gpnoise.h:31): double, returned as float
  const double var = ((double)g.ppc / std::pow((double)g.kr, 3.0)) * std::pow(M_PI, 1.5);
  g.sqrt_var = std::sqrt((float)var);
  g.nk2 = -2.0f / g.L2;
  g.k4 = 4.0f / (g.L2 * g.L2);
  g.bound = std::fabs(g.sigma) * 27.0f * (float)g.ppc / g.sqrt_var * 1.00001f + 1e-6f;
  g.grad_h_max = (float)(std::sqrt(2.0) * std::exp(-0.5) / std::sqrt((double)g.L2) * 1.0001);
  return g;
}

mupsi_status lower_scene(const mupsi_scene_desc& d, HostScene* out, std::string* err) {
  HostScene& H = *out;
  dev::DevScene& S = H.ds;
  if (d.width <= 0 || d.height <= 0 || (int64_t)d.width * d.height > (1LL << 28)) {
    *err = "image size out of range";
    return MUPSI_ERR_INVALID;
  }
  if (d.n_primitives < 0 || (d.n_primitives > 0 && !d.primitives)) {
    *err = "bad primitive list";
    return MUPSI_ERR_INVALID;
  }
  // Camera (camera.h:14-19)
  const V dir = vnormalized(vld(d.cam_dir)), up = vnormalized(vld(d.cam_up));
  const V right = vnormalized(vcross(dir, up));
  const V tup = vnormalized(vcross(right, dir));
  for (int a = 0; a < 3; ++a) S.cam_pos[a] = d.cam_pos[a];
  S.cam_dir[0] = dir.x; S.cam_dir[1] = dir.y; S.cam_dir[2] = dir.z;
  S.cam_right[0] = right.x; S.cam_right[1] = right.y; S.cam_right[2] = right.z;
  S.cam_tup[0] = tup.x; S.cam_tup[1] = tup.y; S.cam_tup[2] = tup.z;
  S.aspect = (float)d.width / (float)d.height;
  S.tan_half = std::tan(d.fov / 2.0f);  // degrees fed to tan() as radians (camera.cpp:10-11)
  S.w = d.width;
  S.h = d.height;
  S.max_bounce = d.max_bounce;
  S.gp_mode = d.gp_mode == MUPSI_GP_RENEWAL_PLUS ? 1 : 0;
  S.skip_medium = d.skip_medium ? 1 : 0;
  S.has_gp = d.has_gp_medium ? 1 : 0;
  S.medium_kind = d.has_gp_medium ? d.medium_kind : MUPSI_MEDIUM_GP;
  if (S.medium_kind < MUPSI_MEDIUM_GP || S.medium_kind > MUPSI_MEDIUM_GRID) {
    *err = "unknown medium kind";
    return MUPSI_ERR_INVALID;
  }
  S.mean_type = d.mean_type;
  for (int a = 0; a < 3; ++a) S.mean_c[a] = d.mean_center[a];
  S.mean_r = d.mean_radius;
  S.eps = 1e-4f;
  S.near_eps = (float)((1e-4) * 15);
  S.single_seed_i = host_constant_sampler_i(42u);
  if (S.has_gp && S.medium_kind == MUPSI_MEDIUM_GRID) {
    // ExternalGPMedium: GridMeanFunction(res, bounds) with inv cell = res / (max - min)
    const mupsi_grid_medium_desc& g = d.grid;
    if (!g.mean || g.resolution < 2 || g.resolution > 1024) {
      *err = "external_gp_medium: mean grid missing or resolution outside [2, 1024]";
      return MUPSI_ERR_INVALID;
    }
    for (int a = 0; a < 3; ++a) {
      if (!(g.bounds_max[a] > g.bounds_min[a]) || !std::isfinite(g.bounds_max[a] - g.bounds_min[a])) {
        *err = "external_gp_medium: bounds_max must exceed bounds_min";
        return MUPSI_ERR_INVALID;
      }
      S.grid.bmin[a] = g.bounds_min[a];
      S.grid.inv[a] = (float)g.resolution / (g.bounds_max[a] - g.bounds_min[a]);
    }
    const size_t n = (size_t)g.resolution * g.resolution * g.resolution;
    H.grid_mean.assign(g.mean, g.mean + n);
    if (g.variance) H.grid_var.assign(g.variance, g.variance + n);
    S.grid.res = g.resolution;
    S.grid.noise_scale = g.noise_scale;
  }
  if (S.has_gp && S.medium_kind == MUPSI_MEDIUM_GP && !validate_gp(d.kernel, err))
    return MUPSI_ERR_INVALID;
  if (S.has_gp && S.medium_kind != MUPSI_MEDIUM_GRID) {
    S.gp = S.medium_kind == MUPSI_MEDIUM_GP ? make_dev_gp(d.kernel) : dev::DevGP{};
    if (S.medium_kind != MUPSI_MEDIUM_GP) {
      S.gp.ppc = 1;
      S.gp.kr = 1.0f;
      S.gp.L2 = 1.0f;
      S.gp.sqrt_var = 1.0f;
    }
    if (d.mean_type == MUPSI_MEAN_MESH) {
      if (!lower_mesh(d.mean_mesh, &H.mean_mesh, err)) return MUPSI_ERR_INVALID;
      H.has_mean_mesh = true;
    } else if (d.mean_type != MUPSI_MEAN_SPHERE) {
      *err = "unknown mean_type";
      return MUPSI_ERR_INVALID;
    }
  } else {
    S.gp = dev::DevGP{};
    S.gp.ppc = 1;
    S.gp.kr = 1.0f;
    S.gp.L2 = 1.0f;
    S.gp.sqrt_var = 1.0f;
  }
  if (texture_ok(d.sky)) {
    H.sky.assign(d.sky.rgb, d.sky.rgb + 3LL * d.sky.width * d.sky.height);
    S.sky.w = d.sky.width;
    S.sky.h = d.sky.height;
  }
  S.n_lights = 0;
  for (int i = 0; i < d.n_primitives; ++i) {
    const mupsi_primitive_desc& pd = d.primitives[i];
    dev::DevPrim pr{};
    pr.type = pd.type;
    pr.bsdf = pd.bsdf;
    pr.mesh = -1;
    if (pd.bsdf < MUPSI_BSDF_LAMBERTIAN || pd.bsdf > MUPSI_BSDF_SPECULAR) {
      *err = "unknown bsdf";
      return MUPSI_ERR_INVALID;
    }
    std::vector<float> albedo;
    if (pd.type == MUPSI_PRIM_SPHERE) {
      if (!(pd.radius != 0.0f) || !std::isfinite(pd.radius)) {
        *err = "sphere radius must be finite and non-zero";
        return MUPSI_ERR_INVALID;
      }
      pr.xf = make_xf(pd.center, pd.radius);
      pr.has_em = pd.has_emission ? 1 : 0;
      for (int a = 0; a < 3; ++a) pr.em[a] = pd.emission[a];
      pr.int_medium = pd.int_medium ? 1 : 0;
      if (pr.has_em) {
        if (S.n_lights >= 16) {
          *err = "more than 16 emissive primitives";
          return MUPSI_ERR_UNSUPPORTED;
        }
        S.lights[S.n_lights++] = i;
      }
    } else if (pd.type == MUPSI_PRIM_MESH) {
      HostMesh hm;
      if (!lower_mesh(pd.mesh, &hm, err)) return MUPSI_ERR_INVALID;
      H.meshes.push_back(std::move(hm));
      pr.mesh = (int)H.meshes.size() - 1;
      // Mesh emission: the reference parses "emission" for spheres only
      // (controller.cpp:225-228); a mesh given one is lit and sampled exactly as
      // its Mesh::getEmission / Mesh::sampleDirect would be (mesh.cpp:252-296).
      pr.has_em = pd.has_emission ? 1 : 0;
      for (int a = 0; a < 3; ++a) pr.em[a] = pd.emission[a];
      if (pr.has_em) {
        if (S.n_lights >= 16) {
          *err = "more than 16 emissive primitives";
          return MUPSI_ERR_UNSUPPORTED;
        }
        S.lights[S.n_lights++] = i;
      }
      if (pd.bsdf == MUPSI_BSDF_LAMBERTIAN && texture_ok(pd.albedo)) {
        albedo.assign(pd.albedo.rgb, pd.albedo.rgb + 3LL * pd.albedo.width * pd.albedo.height);
        pr.albedo.w = pd.albedo.width;
        pr.albedo.h = pd.albedo.height;
      }
    } else {
      *err = "unknown primitive type";
      return MUPSI_ERR_INVALID;
    }
    H.prims.push_back(pr);
    H.prim_albedo.push_back(std::move(albedo));
  }
  S.n_prims = (int)H.prims.size();
  return MUPSI_OK;
}

mupsi_status lower_gp_medium(int32_t mean_type, const float* mean_center, float mean_radius,
                             const mupsi_mesh_desc* mean_mesh, const mupsi_gp_params& kernel,
                             HostScene* io, std::string* err) {
  dev::DevScene& S = io->ds;
  if (!S.has_gp || S.medium_kind != MUPSI_MEDIUM_GP) {
    *err = "scene has no gp_medium to replace";
    return MUPSI_ERR_INVALID;
  }
  if (!validate_gp(kernel, err)) return MUPSI_ERR_INVALID;
  HostMesh mesh;
  if (mean_type == MUPSI_MEAN_MESH) {
    if (!mean_mesh) {
      *err = "mean_type mesh needs a mesh";
      return MUPSI_ERR_INVALID;
    }
    if (!lower_mesh(*mean_mesh, &mesh, err)) return MUPSI_ERR_INVALID;
  } else if (mean_type == MUPSI_MEAN_SPHERE) {
    if (!mean_center) {
      *err = "mean_type sphere needs a centre";
      return MUPSI_ERR_INVALID;
    }
  } else {
    *err = "unknown mean_type";
    return MUPSI_ERR_INVALID;
  }
  S.gp = make_dev_gp(kernel);
  S.mean_type = mean_type;
  if (mean_type == MUPSI_MEAN_SPHERE) {
    for (int a = 0; a < 3; ++a) S.mean_c[a] = mean_center[a];
    S.mean_r = mean_radius;
  }
  io->has_mean_mesh = mean_type == MUPSI_MEAN_MESH;
  io->mean_mesh = std::move(mesh);
  return MUPSI_OK;
}

}  // namespace mupsi
