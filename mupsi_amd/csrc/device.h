// device.h — gfx950 device functions of the GPIS path tracer.
//
// Float expressions follow one fixed operation order (the library is compiled
// with -ffp-contract=off; the only fused multiply-adds are the explicit
// fmaf() calls), chosen to match what GCC -O3 -march=native makes of the
// reference's Eigen expressions:
//   |v|^2 = fmaf(x,x, fmaf(y,y, z*z)),  a.b likewise,  o + d*t = fmaf(d,t,o).
// Integer work (cell index, make_seed, PCG draws, impulse positions and
// weights, BVH traversal order) is bit-exact with the reference algorithm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "host_bvh.h"

namespace mupsi {
namespace dev {

// ------------------------------------------------------------------ vec3
struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 mk(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul(F3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 cmul(F3 a, F3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ F3 divs(F3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ F3 neg(F3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float sq3(F3 v) { return fmaf(v.x, v.x, fmaf(v.y, v.y, v.z * v.z)); }
__device__ __forceinline__ float dot(F3 a, F3 b) {
  return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
}
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
            fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float norm(F3 v) { return sqrtf(sq3(v)); }
__device__ __forceinline__ F3 normalized(F3 v) {
  float z = sq3(v);
  return z > 0.0f ? divs(v, sqrtf(z)) : v;
}
__device__ __forceinline__ F3 madd(F3 d, float t, F3 o) {
  return mk(fmaf(d.x, t, o.x), fmaf(d.y, t, o.y), fmaf(d.z, t, o.z));
}
__device__ __forceinline__ float comp(F3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ float max3(F3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
__device__ __forceinline__ F3 ld3(const float* p, int64_t i) {
  return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

// ------------------------------------------------------------------ RNG  (random.h:13-119)
__device__ __forceinline__ uint32_t hash_combine(uint32_t h, uint32_t v) {
  return h ^ (v + 0x9e3779b9u + (h << 6) + (h >> 2));
}
__device__ __forceinline__ uint32_t make_seed4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  uint32_t h = hash_combine(0u, a);
  h = hash_combine(h, b);
  h = hash_combine(h, c);
  return hash_combine(h, d);
}
__device__ __forceinline__ uint32_t make_seed5(uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                               uint32_t e) {
  return hash_combine(make_seed4(a, b, c, d), e);
}
// PCG-XSH-RR 64/32, multiplier 6364136223846793005, increment 1.
// The output permutation of the state a draw starts from.
__device__ __forceinline__ uint32_t pcg_output(uint64_t old) {
  const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return __builtin_amdgcn_alignbit(xs, xs, rot);  // rotate right by rot
}
__device__ __forceinline__ uint32_t pcg_next(uint64_t& s) {
  const uint64_t old = s;
  s = old * 6364136223846793005ULL + 1ULL;
  return pcg_output(old);
}
// Random::normalizedUint: mantissa fill, [0, 1)
__device__ __forceinline__ float u01(uint32_t i) {
  return __uint_as_float((i >> 9u) | 0x3F800000u) - 1.0f;
}
struct Pcg {
  uint64_t s;
  __device__ __forceinline__ uint32_t nextI() { return pcg_next(s); }
  __device__ __forceinline__ float next1D() { return u01(pcg_next(s)); }
};
// Random(seed): zero seed becomes 1, then two warm-up draws.
__device__ __forceinline__ uint64_t pcg_ctor_state(uint64_t seed) {
  uint64_t s = seed == 0 ? 1ULL : seed;
  pcg_next(s);
  pcg_next(s);
  return s;
}
// ConstantSampler(seed).nextI(): the 5th output of Random(seed).
__device__ __forceinline__ uint32_t constant_sampler_i(uint32_t seed) {
  uint64_t s = pcg_ctor_state(seed);
  pcg_next(s);
  pcg_next(s);
  return pcg_next(s);
}

// ------------------------------------------------------------------ scene structs
struct DevXf {  // Translation(pos) * Scaling(s) and its inverse
  float s, inv_s;
  float pos[3], inv_t[3];
  int identity;
};
struct DevTex {
  const float* rgb;
  int w, h;
};
struct DevMesh {
  // internal nodes as 64-byte records of both children's bounds, breadth-first
  // from the root (record 0); the first ntop may be read from `top`, a copy in
  // LDS that a kernel makes at its start (null where it makes none)
  const BvhRec* recs;
  const BvhRec* top;
  uint32_t ntop, n_recs;
  uint32_t root_ref;         // the root: record 0, or a leaf reference for <= 4 triangles
  uint32_t occ_ref;          // node 0 of the reference's numbering (a leaf): BVH::occluded
  float occ_min[3], occ_max[3];
  // inside() may walk the tree once without a stack: it is at most 60 levels
  // deep, so that the reference's 64-entry stack can never fill
  int stackless;
  const float4* leaf_tris;   // 3 per leaf slot: (v0, id bits), (v1, 0), (v2, 0)
  const float* vnormals;     // per vertex, or null
  const int32_t* faces;
  const float* fnormals;
  const float* texcoords;    // or null
  const int32_t* face_tex;   // or null
  const float* verts;        // local vertex positions (light sampling)
  const float* face_cdf;     // normalised cumulative face areas (light sampling)
  float total_area;          // local-space area (Mesh::prepareForRender)
  int n_nodes, n_faces;      // n_nodes: nodes of the reference's tree
  DevXf xf;
  float lbmin[3], lbmax[3];  // local-space bounds of the whole mesh
  // distance to the mesh from the centres c of a grid of cubes (local space), or
  // null: dist(p, mesh) lies within dlb[(iz * dn[1] + iy) * dn[0] + ix] +- |p - c|
  // (widened by a relative 1e-5 and dslack for rounding) for every p of that
  // cube, the distance being 1-Lipschitz; c = dorg + (i + 1/2) dcell per axis
  const float* dlb;
  int dn[3];
  float dorg[3], dinv, dcell, dslack;
  // a coarser grid of the same kind spanning a larger box around the first, or
  // null: it bounds the distance at points outside the first grid (points the
  // marcher evaluates where the noise and the conditioning can reach the mean
  // surface's value), where the exact search would otherwise start unbounded
  const float* olb;
  int on[3];
  float oorg[3], oinv, ocell, oslack;
  // a third, coarsest grid of the same kind around the second, or null: the far
  // points that conditioning grown over many Renewal+ updates can still bring
  // near zero
  const float* flb;
  int fn[3];
  float forg[3], finv, fcell, fslack;
  // the leaves binned by their y-z bounds (mean meshes whose tree allows the
  // stackless walk), or null: bin (iy, iz) = ((y - xb_org[0]) * xb_inv,
  // (z - xb_org[1]) * xb_inv) truncated, entries xb_off[b] .. xb_off[b + 1] - 1
  // of xb_ent (two float4 each: xmax, ymin, ymax, zmin | zmax, xmin, leaf
  // reference bits, 0), sorted by xmax, largest first
  const float4* xb_ent;
  const uint32_t* xb_off;
  int xb_n[2];
  float xb_org[2], xb_inv;
};
struct DevPrim {
  int type, bsdf;
  DevXf xf;
  int has_em;
  float em[3];
  int int_medium;
  int mesh;
  DevTex albedo;
};
struct DevGP {
  float sigma, kr, L2, nk2, sqrt_var;  // nk2 = -2/L2
  float k4;                            // 4/(L2*L2)
  float bound;                         // |psi| <= bound for every p
  float grad_h_max;                    // max |grad h| = sqrt(2) e^{-1/2} / L
  int ppc;
};
// ExternalGPMedium's grids (GridMeanFunction, meanfunction.cpp:41-86).
struct DevGrid {
  const float* mean;
  const float* var;  // or null
  int res;
  float bmin[3], inv[3];  // inv = res / (bmax - bmin)
  float noise_scale;
};

struct DevScene {
  float cam_pos[3], cam_dir[3], cam_right[3], cam_tup[3];
  float aspect, tan_half;
  int w, h;
  int max_bounce, gp_mode, skip_medium, has_gp;  // has_gp: a medium of medium_kind exists
  int medium_kind;                               // MUPSI_MEDIUM_*
  DevGrid grid;
  int mean_type;
  float mean_c[3], mean_r;
  DevMesh mean_mesh;
  DevGP gp;
  DevTex sky;
  const DevPrim* prims;
  int n_prims;
  const DevMesh* meshes;
  int n_lights;
  int lights[16];
  float eps, near_eps;
  uint32_t single_seed_i;  // ConstantSampler(42).nextI()
};

struct Cond {
  bool active;
  F3 C;
  float u;
  F3 g;
};

struct Ray {
  F3 o, d;
  float nearT, farT;
};
__device__ __forceinline__ Ray make_ray(F3 o, F3 d) {
  return Ray{o, normalized(d), 0.0f, 3.402823466e+38f};
}

// ------------------------------------------------------------------ GP noise
// One cell's impulses are regenerated from its seed (no table): cell seed
// make_seed(cx, cy, cz, seed), state cell_seed + 1 (u32 wrap), per impulse
// draws d0..d3, offset (d2, d1, d0) (GCC's right-to-left argument order),
// weight -1 if d3 < 0.5 else +1, position fmaf(offset, kr, n * kr).
// (gpnoise.cpp:57-74).  d3 < 0.5  <=>  the draw's top bit is clear.

__device__ __forceinline__ void cell_index(const DevGP& g, F3 p, int& cx, int& cy, int& cz) {
  cx = (int)(p.x / g.kr);
  cy = (int)(p.y / g.kr);
  cz = (int)(p.z / g.kr);
}

// One impulse's terms of psi's sums at p: (w h, w (p - s)_x k, w (p - s)_y k,
// w (p - s)_z k) with h = exp(-|p - s|^2 / L^2), k = h (-2 / L^2), s the impulse
// position and w = +-1 its weight (kernel.h:57-66).  Adding w t is adding or
// subtracting t exactly as the reference's `w * h` terms do, since w * t = +-t
// exactly.  psi_eval and the marcher's exact sums at a medium hit share it.
__device__ __forceinline__ float4 psi_terms(const DevGP& g, F3 p, float sx, float sy, float sz,
                                            float w) {
  const float ddx = p.x - sx;
  const float ddy = p.y - sy;
  const float ddz = p.z - sz;
  const float r2v = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
  const float h = expf(-r2v / g.L2);
  const float k = h * g.nk2;
  return make_float4(w * h, w * (ddx * k), w * (ddy * k), w * (ddz * k));
}

// RawNoise's normalisation sigma * sum / sqrt(Var) of the accumulated sums.
__device__ __forceinline__ void psi_finish(const DevGP& g, float sv, float sx, float sy, float sz,
                                           float& v, F3& grad) {
  v = g.sigma * sv / g.sqrt_var;
  grad = mk(g.sigma * sx / g.sqrt_var, g.sigma * sy / g.sqrt_var, g.sigma * sz / g.sqrt_var);
}

// psi value and gradient at p (RawNoise, gpnoise.h:44); sum of |h| in *mass.
// The sums run over the 27 neighbour cells in (dx, dy, dz) order, -1..1 each,
// and over a cell's impulses in draw order: one float addition chain per sum.
template <bool kGrad>
__device__ __forceinline__ void psi_eval(const DevGP& g, F3 p, uint32_t seed, float& v,
                                         F3& grad, float* mass) {
  int cx, cy, cz;
  cell_index(g, p, cx, cy, cz);
  float sv = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f, m = 0.0f;
  for (int dx = -1; dx <= 1; ++dx) {
    const int nx = cx + dx;
    const float bx = (float)nx * g.kr;
    for (int dy = -1; dy <= 1; ++dy) {
      const int ny = cy + dy;
      const float by = (float)ny * g.kr;
      for (int dz = -1; dz <= 1; ++dz) {
        const int nz = cz + dz;
        const float bz = (float)nz * g.kr;
        uint64_t s = (uint32_t)(make_seed4((uint32_t)nx, (uint32_t)ny, (uint32_t)nz, seed) + 1u);
        for (int i = 0; i < g.ppc; ++i) {
          const uint32_t r0 = pcg_next(s);
          const uint32_t r1 = pcg_next(s);
          const uint32_t r2 = pcg_next(s);
          const uint32_t r3 = pcg_next(s);
          const float4 t = psi_terms(g, p, fmaf(u01(r2), g.kr, bx), fmaf(u01(r1), g.kr, by),
                                     fmaf(u01(r0), g.kr, bz), (r3 >> 31) ? 1.0f : -1.0f);
          sv = sv + t.x;
          if (kGrad) {
            sx = sx + t.y;
            sy = sy + t.z;
            sz = sz + t.w;
          }
          if (mass) m += fabsf(t.x);
        }
      }
    }
  }
  F3 gr;
  psi_finish(g, sv, sx, sy, sz, v, gr);
  if (kGrad) grad = gr;
  if (mass) *mass = g.sigma * m / g.sqrt_var;
}

__device__ __forceinline__ float kern_h(const DevGP& g, F3 s, F3 p) {
  return expf(-sq3(sub(p, s)) / g.L2);
}

// ------------------------------------------------------------------ transforms
__device__ __forceinline__ F3 xf_to_local(const DevXf& x, F3 p) {
  return mk(fmaf(x.inv_s, p.x, x.inv_t[0]), fmaf(x.inv_s, p.y, x.inv_t[1]),
            fmaf(x.inv_s, p.z, x.inv_t[2]));
}
__device__ __forceinline__ F3 xf_to_world(const DevXf& x, F3 p) {
  return mk(fmaf(x.s, p.x, x.pos[0]), fmaf(x.s, p.y, x.pos[1]), fmaf(x.s, p.z, x.pos[2]));
}

// ------------------------------------------------------------------ BVH
// AABB::intersect (boudingbox.cpp:6-38) up to its farT clause: the slab entry
// tmin where tmax >= tmin and tmax >= nearT, NaN where the box is missed
// whatever farT is.  The box is hit iff slab_tmin(...) <= farT (NaN compares
// false), so a traversal can test a box when it is first seen and complete the
// test with the farT it has when it takes the box up.
__device__ __forceinline__ float slab_tmin(const float* bmin, const float* bmax, const Ray& r) {
  float tmin = -INFINITY, tmax = INFINITY;
  const float o[3] = {r.o.x, r.o.y, r.o.z};
  const float d[3] = {r.d.x, r.d.y, r.d.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (fabsf(d[a]) > 1e-8f) {
      const float t1 = (bmin[a] - o[a]) / d[a];
      const float t2 = (bmax[a] - o[a]) / d[a];
      const float lo = t2 < t1 ? t2 : t1;  // std::min(t1, t2)
      const float hi = t1 < t2 ? t2 : t1;  // std::max(t1, t2)
      tmin = tmin < lo ? lo : tmin;
      tmax = hi < tmax ? hi : tmax;
    } else if (o[a] < bmin[a] || o[a] > bmax[a]) {
      return __builtin_nanf("");
    }
  }
  return (tmax >= tmin && tmax >= r.nearT) ? tmin : __builtin_nanf("");
}

// Moller-Trumbore (bvh.cpp:82-111)
__device__ __forceinline__ bool ray_tri(const Ray& r, F3 v0, F3 v1, F3 v2, float& t, float& u,
                                        float& v) {
  const F3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  const F3 h = cross(r.d, e2);
  const float a = dot(e1, h);
  if (fabsf(a) < 1e-8f) return false;
  const float f = 1.0f / a;
  const F3 s = sub(r.o, v0);
  u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  const F3 q = cross(s, e1);
  v = f * dot(r.d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  t = f * dot(e2, q);
  return t > r.nearT && t < r.farT;
}

// One record: from global memory (64 bytes, one cache line), or with kTop from
// the LDS copy of the top levels where the kernel made one.
template <bool kTop = false>
__device__ __forceinline__ BvhRec load_rec(const DevMesh& m, uint32_t i) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  using LdsV4 = const V4 __attribute__((address_space(3)))*;
  using GlobalV4 = const V4 __attribute__((address_space(1)))*;
  V4 a, b, c, d;
  if (kTop && i < m.ntop) {  // ds_read from the block's copy
    const LdsV4 p = (LdsV4)(m.top + i);
    a = p[0];
    b = p[1];
    c = p[2];
    d = p[3];
  } else {
    const GlobalV4 p = (GlobalV4)(m.recs + i);
    a = p[0];
    b = p[1];
    c = p[2];
    d = p[3];
  }
  BvhRec R;
  R.lmin[0] = a.x;
  R.lmin[1] = a.y;
  R.lmin[2] = a.z;
  R.lmax[0] = a.w;
  R.lmax[1] = b.x;
  R.lmax[2] = b.y;
  R.l = __float_as_uint(b.z);
  R.order = __float_as_uint(b.w);
  R.rmin[0] = c.x;
  R.rmin[1] = c.y;
  R.rmin[2] = c.z;
  R.rmax[0] = c.w;
  R.rmax[1] = d.x;
  R.rmax[2] = d.y;
  R.r = __float_as_uint(d.z);
  R.next = __float_as_int(d.w);
  return R;
}

// One child of record i (its 32-byte half): bounds, reference, and for the
// right child the record's next.
struct RecHalf {
  float bmin[3], bmax[3];
  uint32_t ref;
  int32_t next;
};
template <bool kTop = false>
__device__ __forceinline__ RecHalf load_half(const DevMesh& m, uint32_t i, bool right) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  using LdsV4 = const V4 __attribute__((address_space(3)))*;
  using GlobalV4 = const V4 __attribute__((address_space(1)))*;
  const uint32_t k = 4 * i + (right ? 2 : 0);
  V4 a, b;
  if (kTop && i < m.ntop) {
    const LdsV4 p = (LdsV4)(m.top) + k;
    a = p[0];
    b = p[1];
  } else {
    const GlobalV4 p = (GlobalV4)(m.recs) + k;
    a = p[0];
    b = p[1];
  }
  RecHalf h;
  h.bmin[0] = a.x;
  h.bmin[1] = a.y;
  h.bmin[2] = a.z;
  h.bmax[0] = a.w;
  h.bmax[1] = b.x;
  h.bmax[2] = b.y;
  h.ref = __float_as_uint(b.z);
  h.next = __float_as_int(b.w);
  return h;
}
__device__ __forceinline__ bool is_leaf(uint32_t ref) { return (ref & kLeafFlag) != 0; }
__device__ __forceinline__ uint32_t leaf_first(uint32_t ref) { return ref & kLeafSlotMask; }
__device__ __forceinline__ uint32_t leaf_count(uint32_t ref) {
  return (ref & ~kLeafFlag) >> kLeafCountShift;
}

__device__ __forceinline__ void leaf_tri(const DevMesh& m, uint32_t slot, F3& v0, F3& v1, F3& v2,
                                         uint32_t& id) {
  const float4 a = m.leaf_tris[3 * slot], b = m.leaf_tris[3 * slot + 1],
               c = m.leaf_tris[3 * slot + 2];
  v0 = mk(a.x, a.y, a.z);
  v1 = mk(b.x, b.y, b.z);
  v2 = mk(c.x, c.y, c.z);
  id = __float_as_uint(a.w);
}

constexpr int kStack = 64;  // the reference's traversal stack (bvh.cpp)

// Traversal cost counters: a node is counted when its bounds are tested, a
// triangle when its intersection or distance test runs.  NoCount compiles to
// nothing; TravCount is used by the cost entry point only.
struct NoCount {
  __device__ __forceinline__ void node() {}
  __device__ __forceinline__ void tri() {}
};
struct TravCount {
  uint32_t nodes = 0, tris = 0;
  __device__ __forceinline__ void node() { ++nodes; }
  __device__ __forceinline__ void tri() { ++tris; }
};

// BVH::intersect (bvh.cpp:197-250): front-to-back stack traversal from the root.
// A node's slab test is made when its parent's record is read (both children
// at once) and completed with the current farT when the node is popped, which
// is the reference's test at the pop; the order, the counts and every decision
// are the reference's.
template <bool kTop = false, class C = NoCount>
__device__ bool bvh_intersect(const DevMesh& m, Ray& r, uint32_t& ti, float& u, float& v,
                              C&& cnt = C()) {
  if (m.n_nodes == 0) return false;
  bool hit = false;
  uint32_t sref[kStack];
  float stmin[kStack];
  int sp = 0;
  sref[sp] = m.root_ref;
  stmin[sp++] = slab_tmin(m.lbmin, m.lbmax, r);
  while (sp > 0) {
    --sp;
    const uint32_t ref = sref[sp];
    const float tmin = stmin[sp];
    cnt.node();
    if (!(tmin <= r.farT)) continue;
    if (is_leaf(ref)) {
      const uint32_t s0 = leaf_first(ref), c = leaf_count(ref);
      for (uint32_t i = s0; i < s0 + c; ++i) {
        cnt.tri();
        F3 v0, v1, v2;
        uint32_t id;
        leaf_tri(m, i, v0, v1, v2, id);
        float t, uu, vv;
        if (ray_tri(r, v0, v1, v2, t, uu, vv)) {
          r.farT = t;
          ti = id;
          u = uu;
          v = vv;
          hit = true;
        }
      }
    } else if (sp <= kStack - 2) {  // the reference's 64 entries; deeper trees stop there
      const BvhRec R = load_rec<kTop>(m, ref);
      const int axis = R.order & 3;
      const bool nearFirst = (comp(r.d, axis) > 0) ? ((R.order >> 2) & 1) : ((R.order >> 3) & 1);
      const float tl = slab_tmin(R.lmin, R.lmax, r), tr = slab_tmin(R.rmin, R.rmax, r);
      // the far child first, the near one on top
      sref[sp] = nearFirst ? R.r : R.l;
      stmin[sp++] = nearFirst ? tr : tl;
      sref[sp] = nearFirst ? R.l : R.r;
      stmin[sp++] = nearFirst ? tl : tr;
    }
  }
  return hit;
}

// BVH::occluded (bvh.cpp:254-300): any hit from node 0.  Node 0 of the
// reference's post-order numbering is the first leaf it built, so the
// traversal tests that leaf's bounds and its triangles only.
__device__ bool bvh_occluded(const DevMesh& m, const Ray& r) {
  if (m.n_nodes == 0) return false;
  if (!(slab_tmin(m.occ_min, m.occ_max, r) <= r.farT)) return false;
  const uint32_t s0 = leaf_first(m.occ_ref), c = leaf_count(m.occ_ref);
  for (uint32_t i = s0; i < s0 + c; ++i) {
    F3 v0, v1, v2;
    uint32_t id;
    leaf_tri(m, i, v0, v1, v2, id);
    float t, uu, vv;
    if (ray_tri(r, v0, v1, v2, t, uu, vv)) return true;
  }
  return false;
}

__device__ __forceinline__ float dist2_aabb(F3 p, const float* bmin, const float* bmax) {
  const float dx = fmaxf(fmaxf(bmin[0] - p.x, 0.0f), p.x - bmax[0]);
  const float dy = fmaxf(fmaxf(bmin[1] - p.y, 0.0f), p.y - bmax[1]);
  const float dz = fmaxf(fmaxf(bmin[2] - p.z, 0.0f), p.z - bmax[2]);
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}
__device__ __forceinline__ float dist2_line(F3 p, F3 a, F3 b, F3& out) {
  const F3 ab = sub(b, a), ap = sub(p, a);
  const float ab2 = sq3(ab);
  if (ab2 < 1e-6f) {
    out = a;
    return sq3(ap);
  }
  float t = dot(ap, ab) / ab2;
  t = fminf(fmaxf(t, 0.0f), 1.0f);
  const F3 proj = madd(ab, t, a);
  out = proj;
  return sq3(sub(p, proj));
}
// Point-triangle squared distance (bvh.cpp:42-80), including the reference's
// |e1 x e2|^2 < 1e-6 "degenerate" test that falls back to vertex a.
__device__ __forceinline__ float dist2_tri(F3 p, F3 a, F3 b, F3 c, F3& out) {
  const F3 e1 = sub(a, b), e2 = sub(a, c);
  F3 nrm = cross(e1, e2);
  if (sq3(nrm) < 1e-6f) {
    out = a;
    return sq3(sub(p, a));
  }
  nrm = normalized(nrm);
  const float dp = dot(sub(p, a), nrm);
  const F3 pj = mk(fmaf(-dp, nrm.x, p.x), fmaf(-dp, nrm.y, p.y), fmaf(-dp, nrm.z, p.z));
  const F3 ab = neg(e1), bc = sub(c, b), ca = e2;
  const F3 c1 = cross(ab, sub(pj, a)), c2 = cross(bc, sub(pj, b)), c3 = cross(ca, sub(pj, c));
  if (dot(c1, nrm) >= 0 && dot(c2, nrm) >= 0 && dot(c3, nrm) >= 0) {
    out = pj;
    return dp * dp;
  }
  float mn = INFINITY, d;
  F3 tmp;
  if ((d = dist2_line(p, a, b, tmp)) < mn) { mn = d; out = tmp; }
  if ((d = dist2_line(p, b, c, tmp)) < mn) { mn = d; out = tmp; }
  if ((d = dist2_line(p, c, a, tmp)) < mn) { mn = d; out = tmp; }
  return mn;
}

// BVH::closest (bvh.cpp:302-353): best-first with |.|^2 pruning.  md0 (default
// FLT_MAX, the reference's start) may be any value strictly above the squared
// distance the search returns: the first triangle in traversal order that attains
// the minimum is still reached (every node above it is nearer than md0 and than
// any distance found before it), so ti, out and the result are unchanged while
// nodes farther than md0 are never visited.
// A node's squared distance is computed when its parent's record is read and
// kept beside it on the stack; the pop compares the kept value with the current
// minimum, as the reference's recomputation at the pop does.
template <bool kTop = false, class C = NoCount>
__device__ float bvh_closest(const DevMesh& m, F3 p, uint32_t& ti, F3& out, C&& cnt = C(),
                             float md0 = 3.402823466e+38f) {
  if (m.n_nodes == 0) return 0.0f;
  uint32_t sref[kStack];
  float sd[kStack];
  int sp = 0;
  sref[sp] = m.root_ref;
  sd[sp++] = dist2_aabb(p, m.lbmin, m.lbmax);
  float md = md0;
  while (sp > 0) {
    --sp;
    const uint32_t ref = sref[sp];
    cnt.node();
    if (sd[sp] >= md) continue;
    if (is_leaf(ref)) {
      const uint32_t s0 = leaf_first(ref), c = leaf_count(ref);
      for (uint32_t i = s0; i < s0 + c; ++i) {
        cnt.tri();
        F3 v0, v1, v2;
        uint32_t id;
        leaf_tri(m, i, v0, v1, v2, id);
        F3 tmp = mk(0, 0, 0);
        const float d = dist2_tri(p, v0, v1, v2, tmp);
        if (d < md) {
          md = d;
          ti = id;
          out = tmp;
        }
      }
    } else {
      const BvhRec R = load_rec<kTop>(m, ref);
      const float dl = dist2_aabb(p, R.lmin, R.lmax);
      const float dr = dist2_aabb(p, R.rmin, R.rmax);
      cnt.node();
      cnt.node();
      if (sp > kStack - 2) continue;
      if (dl < dr) {
        if (dr < md) { sref[sp] = R.r; sd[sp++] = dr; }
        if (dl < md) { sref[sp] = R.l; sd[sp++] = dl; }
      } else {
        if (dl < md) { sref[sp] = R.l; sd[sp++] = dl; }
        if (dr < md) { sref[sp] = R.r; sd[sp++] = dr; }
      }
    }
  }
  return sqrtf(md);
}

// Mesh::intersect (mesh.cpp:185-209): narrows r.farT on hit.
template <bool kTop = false>
__device__ bool mesh_intersect(const DevMesh& m, Ray& ray, uint32_t& ti, float& u, float& v) {
  Ray lr = ray;
  float tcorr = 1.0f;
  if (!m.xf.identity) {
    const F3 lo = xf_to_local(m.xf, ray.o);
    const F3 ld = mul(ray.d, m.xf.inv_s);
    const float len = norm(ld);
    tcorr = norm(mul(divs(ld, len), m.xf.s));
    lr = make_ray(lo, ld);
    lr.nearT = ray.nearT * len;
    lr.farT = ray.farT * len;
  }
  const bool hit = bvh_intersect<kTop>(m, lr, ti, u, v);
  if (hit) ray.farT = lr.farT * tcorr;
  return hit;
}
__device__ bool mesh_occluded(const DevMesh& m, const Ray& ray) {
  if (m.xf.identity) return bvh_occluded(m, ray);
  const F3 lo = xf_to_local(m.xf, ray.o);
  const F3 ld = mul(ray.d, m.xf.inv_s);
  const float len = norm(ld);
  Ray lr = make_ray(lo, ld);
  lr.nearT = ray.nearT * len;
  lr.farT = ray.farT * len;
  return bvh_occluded(m, lr);
}
// Mesh::distToPoint (mesh.cpp:325-339)
// bound_local: an upper bound of the local-space distance (INFINITY: none).
template <bool kTop = false, class C = NoCount>
__device__ float mesh_dist(const DevMesh& m, F3 p, uint32_t& ti, F3& out, C&& cnt = C(),
                           float bound_local = INFINITY) {
  const F3 pl = xf_to_local(m.xf, p);
  F3 ol = mk(0, 0, 0);
  ti = 0;
  // widened so that it stays strictly above the rounded squared distance
  const float b = fmaf(bound_local, 1.0001f, 1e-5f);
  bvh_closest<kTop>(m, pl, ti, ol, cnt, fminf(b * b, 3.402823466e+38f));
  out = xf_to_world(m.xf, ol);
  return norm(sub(out, p));
}
// Mesh::inside (mesh.cpp:341-362): parity of +x crossings, restarting 1e-4 past each hit.
//
// The reference casts the +x ray once per crossing, each cast from 1e-4 past the
// previous hit.  One traversal collecting every hit gives the same count when
// (a) each hit t lies in its leaf's x-slab [bmin.x - p.x, bmax.x - p.x] as the
// slab test computes it, so no cast culls the leaf of the hit it must find (its
// ancestors' slabs contain the leaf's); (b) no two hits are within 2e-4, so each
// cast's nearT = h + 1e-4 lies strictly between consecutive hits and the cast
// finds exactly the next one; (c) no cast's 64-entry stack fills, which holds
// for trees of at most 60 levels (m.stackless: the near/far traversal keeps at
// most one waiting sibling per level).  The single traversal is a left-first
// pre-order walk without a stack: at a hit internal node it reads the node's
// record and tests the left child, then the right one, and after a subtree it
// continues with the right child of the record the subtree's `next` names.  It
// tests the boxes the reference's casts would test first (both children of
// every hit internal node).  Otherwise the reference's loop of casts runs as
// written.
//
// Where the mesh carries leaf bins (xb_off), the leaves the walk would reach are
// found without walking: a leaf's triangles are tested by the walk iff the leaf
// passes its own slab test (its ancestors' boxes contain its box, so their
// tests pass whenever its test does: the slab test's x range only widens and
// its y-z conditions only loosen), and a leaf that passes holds pl.y and pl.z
// within its y-z bounds, so it is listed in pl's bin.  The bin's entries are
// slab-tested as the walk tests the leaf, the same triangles are tested, and
// the hits, the conditions (a)-(c) and the parity are the walk's; entries are
// sorted by xmax, so the scan stops at the first leaf whose tmax = xmax - pl.x
// falls below nearT (every later one fails the slab test there as well).
// kWalk: 0 the reference's loop of casts only (its traversal cost), 1 the
// stackless walk, 2 the leaf bins where the mesh has them, else the walk.
template <int kWalk = 2, bool kTop = false, class C = NoCount>
__device__ bool mesh_inside(const DevMesh& m, F3 p, C&& cnt = C()) {
  const F3 pl = xf_to_local(m.xf, p);
  if (m.n_nodes == 0) return false;
  if constexpr (kWalk > 0) {
    if (m.stackless) {
      const Ray r = make_ray(pl, mk(1, 0, 0));
      constexpr int kHits = 8;
      float hs[kHits];
      int nh = 0;
      bool ok = true;
      const auto leaf = [&](uint32_t ref, float bminx, float bmaxx) {
        const float lo = bminx - pl.x, hi = bmaxx - pl.x;  // the slab test's x range
        const uint32_t s0 = leaf_first(ref), c = leaf_count(ref);
        for (uint32_t i = s0; i < s0 + c; ++i) {
          cnt.tri();
          F3 v0, v1, v2;
          uint32_t id;
          leaf_tri(m, i, v0, v1, v2, id);
          float t, uu, vv;
          if (ray_tri(r, v0, v1, v2, t, uu, vv)) {
            ok = ok && t >= lo && t <= hi && nh < kHits;
#pragma unroll
            for (int j = 0; j < kHits; ++j) {
              if (j < nh && fabsf(t - hs[j]) <= 2e-4f + 1e-6f * fabsf(t)) ok = false;
              if (j == nh) hs[j] = t;
            }
            ++nh;
          }
        }
      };
      cnt.node();  // the root's box, or a bin's offsets
      if (kWalk == 2 && m.xb_off) {
        const float fy = (pl.y - m.xb_org[0]) * m.xb_inv, fz = (pl.z - m.xb_org[1]) * m.xb_inv;
        // outside every bin, pl lies outside every leaf's y-z bounds: no hit
        if (fy >= 0.0f && fz >= 0.0f && fy < (float)m.xb_n[0] && fz < (float)m.xb_n[1]) {
          const uint32_t b = (uint32_t)fz * (uint32_t)m.xb_n[0] + (uint32_t)fy;
          const uint32_t e1 = m.xb_off[b + 1];
          for (uint32_t e = m.xb_off[b]; e < e1 && ok; ++e) {
            cnt.node();
            const float4 A = m.xb_ent[2 * e], B = m.xb_ent[2 * e + 1];
            if (!(A.x - pl.x >= r.nearT)) break;
            const float bmin[3] = {B.y, A.y, A.w}, bmax[3] = {A.x, A.z, B.x};
            if (slab_tmin(bmin, bmax, r) <= r.farT) leaf(__float_as_uint(B.z), B.y, A.x);
          }
        }
      } else if (slab_tmin(m.lbmin, m.lbmax, r) <= r.farT) {
        if (is_leaf(m.root_ref)) {
          leaf(m.root_ref, m.lbmin[0], m.lbmax[0]);
        } else {
          // the current child: record rec, left or right half
          uint32_t rec = 0;
          bool right = false;
          RecHalf H = load_half<kTop>(m, rec, right);
          uint32_t steps = 0;  // the links are checked on the host; a bound all the same
          while (ok) {
            cnt.node();
            const bool h = slab_tmin(H.bmin, H.bmax, r) <= r.farT;
            if (++steps > 2 * m.n_recs + 2) {
              ok = false;
            } else if (h && !is_leaf(H.ref)) {  // descend to the left child
              rec = H.ref;
              right = false;
              H = load_half<kTop>(m, rec, false);
            } else {
              if (h) leaf(H.ref, H.bmin[0], H.bmax[0]);
              if (!right) {  // the right child of the same record
                right = true;
                H = load_half<kTop>(m, rec, true);
              } else if (H.next < 0) {
                break;
              } else {  // the right child of the record the subtree continues with
                rec = (uint32_t)H.next;
                H = load_half<kTop>(m, rec, true);
              }
            }
          }
        }
      }
      if (ok) return (nh & 1) == 1;
    }
  }
  Ray r = make_ray(pl, mk(1, 0, 0));
  int hits = 0;
  uint32_t ti;
  float u, v;
  if (bvh_intersect<kTop>(m, r, ti, u, v, cnt)) {
    ++hits;
    while (r.farT < INFINITY) {
      r.nearT = r.farT + 1e-4f;
      r.farT = INFINITY;
      if (!bvh_intersect<kTop>(m, r, ti, u, v, cnt)) break;
      ++hits;
    }
  }
  return (hits % 2) == 1;
}
// A distance grid's bounds at pl: the exact distance dc from the centre c of
// pl's cube, widened by |pl - c| (distance is 1-Lipschitz) and for rounding.
// False where pl lies outside the grid.
__device__ __forceinline__ bool dist_grid_at(const float* g, const int* n, const float* org, float inv,
                                             float cell, float slack, F3 pl, float& lo, float& hi) {
  const float fx = (pl.x - org[0]) * inv, fy = (pl.y - org[1]) * inv, fz = (pl.z - org[2]) * inv;
  if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)n[0] && fy < (float)n[1] &&
        fz < (float)n[2]))
    return false;
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const float dc = g[((size_t)iz * n[1] + iy) * n[0] + ix];
  // the cube's centre as the grid was built from it
  const F3 c = mk(org[0] + (ix + 0.5f) * cell, org[1] + (iy + 0.5f) * cell, org[2] + (iz + 0.5f) * cell);
  const float r = sqrtf(sq3(sub(pl, c))) * (1.0f + 1e-5f) + slack;
  lo = fmaxf(lo, dc - r);
  hi = dc + r;
  return true;
}
// Bounds of the local-space distance from pl to the mesh: lo from the mesh's
// bounding box and the distance grids, hi from the finest grid that holds pl
// (infinite outside all three).
__device__ __forceinline__ void mesh_dist_bounds_local(const DevMesh& m, F3 pl, float& lo,
                                                       float& hi) {
  const float dx = fmaxf(fmaxf(m.lbmin[0] - pl.x, 0.0f), pl.x - m.lbmax[0]);
  const float dy = fmaxf(fmaxf(m.lbmin[1] - pl.y, 0.0f), pl.y - m.lbmax[1]);
  const float dz = fmaxf(fmaxf(m.lbmin[2] - pl.z, 0.0f), pl.z - m.lbmax[2]);
  lo = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
  hi = INFINITY;
  if (m.dlb && dist_grid_at(m.dlb, m.dn, m.dorg, m.dinv, m.dcell, m.dslack, pl, lo, hi)) return;
  if (m.olb && dist_grid_at(m.olb, m.on, m.oorg, m.oinv, m.ocell, m.oslack, pl, lo, hi)) return;
  if (m.flb) (void)dist_grid_at(m.flb, m.fn, m.forg, m.finv, m.fcell, m.fslack, pl, lo, hi);
}
// Whether the +x parity ray of Mesh::inside can reach the mesh's bounding box
// from pl (otherwise inside() is exactly false).
__device__ __forceinline__ bool mesh_in_parity_slab(const DevMesh& m, F3 pl) {
  return pl.y >= m.lbmin[1] && pl.y <= m.lbmax[1] && pl.z >= m.lbmin[2] && pl.z <= m.lbmax[2] &&
         m.lbmax[0] - pl.x >= 0.0f;
}

// Mesh::distToPoint with the closest-point search started from the distance
// grid's upper bound at p (same ti, closest point and distance; fewer nodes).
template <bool kTop = false, class C = NoCount>
__device__ float mesh_dist_grid(const DevMesh& m, F3 p, uint32_t& ti, F3& out, C&& cnt = C()) {
  float lo, hi;
  mesh_dist_bounds_local(m, xf_to_local(m.xf, p), lo, hi);
  return mesh_dist<kTop>(m, p, ti, out, cnt, hi);
}

__device__ __forceinline__ F3 mesh_world_normal(const DevMesh& m, F3 n) {
  if (m.xf.identity) return n;
  return normalized(mul(n, m.xf.inv_s));
}

// ------------------------------------------------------------------ mean + field
// The mesh mean (meanfunction.cpp:22-39): the sign from Mesh::inside, the
// distance from Mesh::distToPoint (searched from the distance grids' bound), the
// gradient along p - closest point (the face normal on the surface).  kTop: the
// closest-point traversal reads the top records from the kernel's LDS copy.
template <bool kGrad, bool kTop = false>
__device__ __forceinline__ float mesh_mean_eval(const DevMesh& M, F3 p, F3* grad) {
  const float sg = mesh_inside(M, p) ? -1.0f : 1.0f;
  uint32_t ti = 0;
  F3 o;
  const float d = mesh_dist_grid<kTop>(M, p, ti, o);
  if (kGrad) {
    if (d < 1e-6f) {
      const float* fnp = M.fnormals + 3 * ti;
      *grad = mesh_world_normal(M, mk(fnp[0], fnp[1], fnp[2]));
    } else {
      *grad = normalized(mul(sub(p, o), sg));
    }
  }
  return sg * d;
}

// mu and grad mu (meanfunction.cpp:13-39).  For a mesh mean the reference calls
// inside() and distToPoint() once for the value and again for the gradient;
// both calls return the same bits, so one pair serves both here.
// kMesh false: the mean is known to be a sphere (no BVH code is compiled).
template <bool kGrad, bool kMesh = true>
__device__ __forceinline__ float mean_eval(const DevScene& S, F3 p, F3* grad) {
  if (!kMesh || S.mean_type == 0) {
    const F3 r = sub(p, mk(S.mean_c[0], S.mean_c[1], S.mean_c[2]));
    if (kGrad) *grad = normalized(r);
    return norm(r) - S.mean_r;
  }
  return mesh_mean_eval<kGrad>(S.mean_mesh, p, grad);
}

// The mean and its gradient at one point, for callers that need them twice.
struct MeanAt {
  float mu;
  F3 grad;
};

// ------------------------------------------------------------------ SDF / grid media
// GridMeanFunction::eval (meanfunction.cpp:56-76): trilinear over cell centres,
// indices clamped to [0, res - 2], weights clamped to [0, 1]; data x-major.
__device__ __forceinline__ float grid_eval(const DevGrid& g, const float* data, F3 p) {
  const float fx = fmaf(p.x - g.bmin[0], g.inv[0], -0.5f);
  const float fy = fmaf(p.y - g.bmin[1], g.inv[1], -0.5f);
  const float fz = fmaf(p.z - g.bmin[2], g.inv[2], -0.5f);
  const int i0 = max(0, min(g.res - 2, (int)floorf(fx)));
  const int j0 = max(0, min(g.res - 2, (int)floorf(fy)));
  const int k0 = max(0, min(g.res - 2, (int)floorf(fz)));
  const float wx = fminf(fmaxf(fx - (float)i0, 0.0f), 1.0f);
  const float wy = fminf(fmaxf(fy - (float)j0, 0.0f), 1.0f);
  const float wz = fminf(fmaxf(fz - (float)k0, 0.0f), 1.0f);
  const size_t sx = (size_t)g.res * g.res, sy = (size_t)g.res;
  const auto at = [&](int i, int j, int k) { return data[i * sx + j * sy + k]; };
  const float ux = 1.0f - wx, uy = 1.0f - wy, uz = 1.0f - wz;
  const float c00 = fmaf(at(i0 + 1, j0, k0), wx, at(i0, j0, k0) * ux);
  const float c01 = fmaf(at(i0 + 1, j0, k0 + 1), wx, at(i0, j0, k0 + 1) * ux);
  const float c10 = fmaf(at(i0 + 1, j0 + 1, k0), wx, at(i0, j0 + 1, k0) * ux);
  const float c11 = fmaf(at(i0 + 1, j0 + 1, k0 + 1), wx, at(i0, j0 + 1, k0 + 1) * ux);
  const float c0 = fmaf(c10, wy, c00 * uy);
  const float c1 = fmaf(c11, wy, c01 * uy);
  return fmaf(c1, wz, c0 * uz);
}
// GridMeanFunction::gradient (meanfunction.cpp:78-84): central differences, eps 1e-4.
__device__ __forceinline__ F3 grid_grad(const DevGrid& g, const float* data, F3 p) {
  const float e = 1e-4f, e2 = 2.0f * e;
  const float dx = grid_eval(g, data, mk(p.x + e, p.y, p.z)) - grid_eval(g, data, mk(p.x - e, p.y, p.z));
  const float dy = grid_eval(g, data, mk(p.x, p.y + e, p.z)) - grid_eval(g, data, mk(p.x, p.y - e, p.z));
  const float dz = grid_eval(g, data, mk(p.x, p.y, p.z + e)) - grid_eval(g, data, mk(p.x, p.y, p.z - e));
  return mk(dx / e2, dy / e2, dz / e2);
}
// hash11 (externalgpmedium.cpp:8-17): the float bits of p mixed into seed.
__device__ __forceinline__ float hash11(float x, float y, float z, uint32_t seed) {
  uint32_t h = seed;
  h = h * 0x9e3779b1u + __float_as_uint(x);
  h = h * 0x9e3779b1u + __float_as_uint(y);
  h = h * 0x9e3779b1u + __float_as_uint(z);
  h = (h ^ (h >> 16)) * 0x85ebca6bu;
  h = (h ^ (h >> 13)) * 0xc2b2ae35u;
  h = h ^ (h >> 16);
  return (float)(h & 0xFFFFFFu) / 16777215.0f * 2.0f - 1.0f;
}
// ExternalGPMedium::evalWithGradient (externalgpmedium.cpp:31-58): grid mean
// plus noise_scale * sqrt(variance) * hash(p, seed); the gradient adds the
// hash's central differences (eps 0.01).
template <bool kGrad>
__device__ __forceinline__ float grid_medium_eval(const DevScene& S, F3 p, uint32_t seed, F3* grad) {
  const DevGrid& g = S.grid;
  const float mu = grid_eval(g, g.mean, p);
  float noise = 0.0f;
  F3 ng = mk(0, 0, 0);
  if (g.var && g.noise_scale > 0.0f) {
    const float var = grid_eval(g, g.var, p);
    if (var > 0.0f) {
      const float amp = g.noise_scale * sqrtf(var);
      noise = amp * hash11(p.x, p.y, p.z, seed);
      if (kGrad) {
        const float e = 0.01f, e2 = 2.0f * e;
        ng.x = amp * (hash11(p.x + e, p.y, p.z, seed) - hash11(p.x - e, p.y, p.z, seed)) / e2;
        ng.y = amp * (hash11(p.x, p.y + e, p.z, seed) - hash11(p.x, p.y - e, p.z, seed)) / e2;
        ng.z = amp * (hash11(p.x, p.y, p.z + e, seed) - hash11(p.x, p.y, p.z - e, seed)) / e2;
      }
    }
  }
  if (kGrad) {
    const F3 mg = grid_grad(g, g.mean, p);
    *grad = mk(mg.x + ng.x, mg.y + ng.y, mg.z + ng.z);
  }
  return mu + noise;
}

// GPMedium::evalWithGradient's sum (gpmedium.cpp:19-41) of the mean (mu, grad mu),
// psi (pv, grad psi) and the conditioning term of state c, at p.
template <bool kGrad>
__device__ __forceinline__ float field_combine(const DevGP& gp, F3 p, const Cond& c, float mu,
                                               F3 mg, float pv, F3 pg, F3* grad) {
  float f = mu + pv;
  F3 g = mk(mg.x + pg.x, mg.y + pg.y, mg.z + pg.z);
  if (c.active) {
    const F3 r = sub(p, c.C);
    const float h = kern_h(gp, c.C, p);
    const F3 hg = mul(r, h * gp.nk2);
    f = f + fmaf(c.u, h, dot(c.g, hg));
    if (kGrad) {
      const float k4 = gp.k4 * dot(r, c.g);
      const F3 in = mk(fmaf(k4, r.x, gp.nk2 * c.g.x), fmaf(k4, r.y, gp.nk2 * c.g.y),
                       fmaf(k4, r.z, gp.nk2 * c.g.z));
      const F3 Hg = mul(in, h);
      g = mk(g.x + fmaf(c.u, hg.x, Hg.x), g.y + fmaf(c.u, hg.y, Hg.y), g.z + fmaf(c.u, hg.z, Hg.z));
    }
  }
  if (kGrad) *grad = g;
  return f;
}

// GPMedium::evalWithGradient(p, seed, state) (gpmedium.cpp:19-41).  mean: the
// mean's value and gradient at p when the caller has them already.
template <bool kGrad, bool kMesh = true>
__device__ __forceinline__ float field_eval(const DevScene& S, F3 p, uint32_t seed, const Cond& c,
                                            F3* grad, const MeanAt* mean = nullptr) {
  F3 mg = mk(0, 0, 0), pg = mk(0, 0, 0);
  float mu;
  if (mean) {
    mu = mean->mu;
    mg = mean->grad;
  } else {
    mu = mean_eval<kGrad, kMesh>(S, p, &mg);
  }
  float pv;
  psi_eval<kGrad>(S.gp, p, seed, pv, pg, nullptr);
  return field_combine<kGrad>(S.gp, p, c, mu, mg, pv, pg, grad);
}

// Bound on |conditioning term| for every p: |u| h + |g| |grad h| <= |u| + |g| max|grad h|.
__device__ __forceinline__ float cond_bound(const DevGP& g, const Cond& c) {
  if (!c.active) return 0.0f;
  return (fabsf(c.u) + norm(c.g) * g.grad_h_max) * 1.0001f + 1e-6f;
}

}  // namespace dev
}  // namespace mupsi
