// oracle.cpp — TEST INFRASTRUCTURE ONLY (the checker, never the product path).
//
// Eigen-free CPU restatement of the mupsi reference hot path.  Each function
// cites the reference file:line it follows.  Compiled with -ffp-contract=off so
// the only fused multiply-adds are the explicit fmaf() calls below; those fix
// one documented op sequence for every expression the reference leaves to the
// compiler (GCC -O3 -march=native contracts a+b*c):
//   sq3(v)      = fmaf(x,x, fmaf(y,y, z*z))          Eigen squaredNorm redux x0+(x1+x2)
//   dot(a,b)    = fmaf(ax,bx, fmaf(ay,by, az*bz))
//   cross(a,b)  = fmaf(ay,bz, -(az*by)), ...
//   o + d*t     = fmaf(d, t, o)   (ray points, impulse positions n*kr + u*kr)
//   a*b+c*d+e*f = fmaf(e,f, fmaf(c,d, a*b))          left-associated scalar sums
//
// Pinning (what checks each part; see tests/):
//   make_seed, PCG, next1D, Bernoulli, ConstantSampler ...... golden vectors from the
//       reference's own random.h compiled here (oracle/_ref) + Python-int restatement
//   cell index, impulse draw order and positions ............. closed forms + KATs
//   psi / grad psi ........................................... FP64 re-evaluation,
//       central differences, statistical variance, single-impulse closed form
//   field with conditioning, Renewal+ ........................ invariants grad f(C)=n,
//       f(C)=mu(C)(1-1/sigma)
//   march .................................................... brute-force t-grid scan
//   BVH build/intersect/occluded/closest/inside .............. brute force over triangles
//   camera, sky lookup, BSDF, sphere ......................... closed forms
//   trace / render ........................................... furnace / sky-only scenes
//   SDF / grid media (sdfmedium.cpp, externalgpmedium.cpp) ... analytic sphere hits,
//       linear fields through the trilinear grid, hash11 vs an integer restatement
// Exact float bits of Eigen-mediated expressions in the reference binary are
// parity unpinned (the reference cannot be built here: Eigen is absent).

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- vec3
struct V3 {
  float x, y, z;
};
static inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mul(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline V3 cmul(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 divs(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float sq3(V3 v) { return fmaf(v.x, v.x, fmaf(v.y, v.y, v.z * v.z)); }
static inline float dot(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
static inline V3 cross(V3 a, V3 b) {
  return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
            fmaf(a.x, b.y, -(a.y * b.x)));
}
static inline float norm(V3 v) { return sqrtf(sq3(v)); }
static inline V3 normalized(V3 v) {  // Eigen normalized(): v / sqrt(|v|^2) if |v|^2 > 0
  float z = sq3(v);
  return z > 0.0f ? divs(v, sqrtf(z)) : v;
}
static inline V3 madd(V3 d, float t, V3 o) {  // o + d*t
  return mk(fmaf(d.x, t, o.x), fmaf(d.y, t, o.y), fmaf(d.z, t, o.z));
}
static inline float comp(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
static inline float max3(V3 v) { return std::max(v.x, std::max(v.y, v.z)); }
static inline float fdiv_pi(float x) { return (float)((double)x / M_PI); }

struct D3 {
  double x, y, z;
};

// ---------------------------------------------------------------- RNG  (random.h:13-119)
static inline uint32_t hash_combine(uint32_t h, uint32_t v) {
  return h ^ (v + 0x9e3779b9u + (h << 6) + (h >> 2));
}
static inline uint32_t make_seed4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  uint32_t h = 0;
  h = hash_combine(h, a);
  h = hash_combine(h, b);
  h = hash_combine(h, c);
  h = hash_combine(h, d);
  return h;
}
static inline uint32_t make_seed5(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e) {
  return hash_combine(make_seed4(a, b, c, d), e);
}

struct Pcg {  // Random
  uint64_t s;
  uint32_t nextI() {  // random.h:94-100
    uint64_t old = s;
    s = old * 6364136223846793005ULL + 1ULL;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((uint32_t)(-(int32_t)rot) & 31u));
  }
  float next1D() { return norm_uint(nextI()); }
  static float norm_uint(uint32_t i) {  // random.h:90-92
    uint32_t b = (i >> 9u) | 0x3F800000u;
    float f;
    std::memcpy(&f, &b, 4);
    return f - 1.0f;
  }
};
static inline Pcg pcg_ctor(uint64_t seed) {  // Random(seed): random.h:77-80
  Pcg r{seed == 0 ? 1ULL : seed};
  r.nextI();
  r.nextI();
  return r;
}
static inline uint32_t constant_sampler_i(uint32_t seed) {  // sampler.h:49-54
  Pcg r = pcg_ctor(seed);
  r.next1D();
  r.next1D();
  return r.nextI();
}

// ---------------------------------------------------------------- GP params
struct GP {
  float sigma, kr, L2, sqrt_var, var;
  int ppc;
};
static GP make_gp(const mupsi_gp_params& p) {
  GP g;
  g.sigma = p.sigma;
  g.kr = p.kernel_radius;
  g.L2 = sq3(mk(p.aniso[0], p.aniso[1], p.aniso[2]));  // lengthScale_.squaredNorm()
  g.ppc = p.points_per_cell;
  // Var() = (impulseDensity_ / pow(kr, 3)) * pow(M_PI, 1.5) in double, returned as float
  double v = ((double)g.ppc / std::pow((double)g.kr, 3.0)) * std::pow(M_PI, 1.5);
  g.var = (float)v;
  g.sqrt_var = sqrtf(g.var);
  return g;
}

// One impulse of a cell: offset draw order (d2, d1, d0) = GCC right-to-left
// evaluation of Vector3f(next1D(), next1D(), next1D()) (gpnoise.cpp:64-74).
struct Impulse {
  V3 s;
  float w;
};
static inline void cell_impulses(const GP& g, int cx, int cy, int cz, uint32_t seed,
                                 uint32_t* cell_seed_out, Impulse* out) {
  uint32_t cs = make_seed4((uint32_t)cx, (uint32_t)cy, (uint32_t)cz, seed);
  if (cell_seed_out) *cell_seed_out = cs;
  Pcg r{(uint64_t)(uint32_t)(cs + 1u)};  // setState(cell_seed + 1u), no warm-up
  V3 base = mk((float)cx * g.kr, (float)cy * g.kr, (float)cz * g.kr);
  for (int i = 0; i < g.ppc; ++i) {
    float d0 = r.next1D();
    float d1 = r.next1D();
    float d2 = r.next1D();
    V3 off = mk(d2, d1, d0);
    float wv = r.next1D() < 0.5f ? -1.0f : 1.0f;
    out[i].s = mk(fmaf(off.x, g.kr, base.x), fmaf(off.y, g.kr, base.y), fmaf(off.z, g.kr, base.z));
    out[i].w = wv;
  }
}

static inline void cell_of(const GP& g, V3 p, int* c) {  // (p / kr).cast<int>()
  c[0] = (int)(p.x / g.kr);
  c[1] = (int)(p.y / g.kr);
  c[2] = (int)(p.z / g.kr);
}

// SparseSEKernel::h  (kernel.h:52-54):  exp(-|p-s|^2 / L^2)
static inline float kern_h(const GP& g, V3 s, V3 p) { return expf(-sq3(sub(p, s)) / g.L2); }
// h_grad (kernel.h:55-58): h * (-2/L^2) * (p - C)
static inline V3 kern_hgrad(const GP& g, V3 c, V3 p) {
  float k = kern_h(g, c, p) * (-2.0f / g.L2);
  return mul(sub(p, c), k);
}
// h_hessian_vec (kernel.h:60-65)
static inline V3 kern_hess_vec(const GP& g, V3 c, V3 p, V3 v) {
  V3 r = sub(p, c);
  float hv = kern_h(g, c, p);
  float k2 = -2.0f / g.L2;
  float k4 = (4.0f / (g.L2 * g.L2)) * dot(r, v);
  V3 in = mk(fmaf(k4, r.x, k2 * v.x), fmaf(k4, r.y, k2 * v.y), fmaf(k4, r.z, k2 * v.z));
  return mul(in, hv);
}

struct F4 {
  float v, gx, gy, gz;
};

// RawNoise (gpnoise.h:44) = sigma * InternalNoise / sqrt(Var)  (gpnoise.cpp:51-87)
static F4 raw_noise_f32(const GP& g, V3 p, uint32_t seed, float* l1) {
  int c[3];
  cell_of(g, p, c);
  float sx = 0, sy = 0, sz = 0, sw = 0, mass = 0;
  Impulse imp[64];
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        cell_impulses(g, c[0] + dx, c[1] + dy, c[2] + dz, seed, nullptr, imp);
        for (int i = 0; i < g.ppc; ++i) {
          float h = kern_h(g, imp[i].s, p);
          sx = fmaf(imp[i].w, h, sx);
          V3 gr = kern_hgrad(g, imp[i].s, p);
          sy = fmaf(imp[i].w, gr.x, sy);
          sz = fmaf(imp[i].w, gr.y, sz);
          sw = fmaf(imp[i].w, gr.z, sw);
          mass += h;
        }
      }
  if (l1) *l1 = g.sigma * mass / g.sqrt_var;
  return F4{g.sigma * sx / g.sqrt_var, g.sigma * sy / g.sqrt_var, g.sigma * sz / g.sqrt_var,
            g.sigma * sw / g.sqrt_var};
}

struct F4d {
  double v, gx, gy, gz;
};
static F4d raw_noise_f64(const GP& g, V3 p, uint32_t seed, double* l1) {
  int c[3];
  cell_of(g, p, c);
  double sx = 0, sy = 0, sz = 0, sw = 0, mass = 0;
  double L2 = g.L2;
  Impulse imp[64];
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        cell_impulses(g, c[0] + dx, c[1] + dy, c[2] + dz, seed, nullptr, imp);
        for (int i = 0; i < g.ppc; ++i) {
          double rx = (double)p.x - imp[i].s.x, ry = (double)p.y - imp[i].s.y,
                 rz = (double)p.z - imp[i].s.z;
          double h = std::exp(-(rx * rx + ry * ry + rz * rz) / L2);
          double k = h * (-2.0 / L2) * imp[i].w;
          sx += imp[i].w * h;
          sy += k * rx;
          sz += k * ry;
          sw += k * rz;
          mass += h;
        }
      }
  double sc = (double)g.sigma / std::sqrt((double)g.var);
  if (l1) *l1 = sc * mass;
  return F4d{sc * sx, sc * sy, sc * sz, sc * sw};
}

// ---------------------------------------------------------------- affine T*S (uniform scale)
// Affine3f(Translation3f(pos) * Scaling(s)).matrix() and its inverse.  The
// inverse entries are fixed as inv_s = 1/s, inv_t = -pos * inv_s (Eigen's 4x4
// inverse rounding is parity unpinned).
struct Xf {
  float s, inv_s;
  V3 pos, inv_t;
  bool identity;  // Eigen isIdentity() (prec 1e-5) of the inverse matrix
};
static Xf make_xf(V3 pos, float s) {
  Xf x;
  x.s = s;
  x.pos = pos;
  x.inv_s = 1.0f / s;
  x.inv_t = mk(-(pos.x * x.inv_s), -(pos.y * x.inv_s), -(pos.z * x.inv_s));
  const float prec = 1e-5f;
  bool diag = std::fabs(x.inv_s - 1.0f) <= prec * std::min(std::fabs(x.inv_s), 1.0f);
  bool off = std::fabs(x.inv_t.x) <= prec && std::fabs(x.inv_t.y) <= prec &&
             std::fabs(x.inv_t.z) <= prec;
  x.identity = diag && off;
  return x;
}
static inline V3 xf_to_local(const Xf& x, V3 p) {
  return mk(fmaf(x.inv_s, p.x, x.inv_t.x), fmaf(x.inv_s, p.y, x.inv_t.y),
            fmaf(x.inv_s, p.z, x.inv_t.z));
}
static inline V3 xf_dir_to_local(const Xf& x, V3 d) { return mul(d, x.inv_s); }
static inline V3 xf_to_world(const Xf& x, V3 p) {
  return mk(fmaf(x.s, p.x, x.pos.x), fmaf(x.s, p.y, x.pos.y), fmaf(x.s, p.z, x.pos.z));
}
static inline V3 xf_dir_to_world(const Xf& x, V3 d) { return mul(d, x.s); }

// ---------------------------------------------------------------- ray  (ray.h:9-29)
struct Ray {
  V3 o, d;
  float nearT, farT;
};
static inline Ray make_ray(V3 o, V3 d) {
  return Ray{o, normalized(d), 0.0f, std::numeric_limits<float>::max()};
}

// ---------------------------------------------------------------- AABB / BVH
struct Node {
  float bmin[3], bmax[3];
  uint32_t left, right;
};
static const uint32_t LEAF = 0x80000000u;

static inline int longest_axis(const Node& n) {  // boudingbox.h longestAxis
  float dx = n.bmax[0] - n.bmin[0], dy = n.bmax[1] - n.bmin[1], dz = n.bmax[2] - n.bmin[2];
  if (dx > dy && dx > dz) return 0;
  if (dy > dz) return 1;
  return 2;
}
static inline float centroid(const Node& n, int a) { return (n.bmin[a] + n.bmax[a]) * 0.5f; }

// AABB::intersect (boudingbox.cpp:6-38)
static bool aabb_hit(const Node& n, const Ray& r) {
  float tmin = -INFINITY, tmax = INFINITY;
  const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
  for (int a = 0; a < 3; ++a) {
    if (std::fabs(d[a]) > 1e-8f) {
      float t1 = (n.bmin[a] - o[a]) / d[a];
      float t2 = (n.bmax[a] - o[a]) / d[a];
      tmin = std::max(tmin, std::min(t1, t2));
      tmax = std::min(tmax, std::max(t1, t2));
    } else if (o[a] < n.bmin[a] || o[a] > n.bmax[a]) {
      return false;
    }
  }
  return tmax >= tmin && tmax >= r.nearT && tmin <= r.farT;
}

struct Tri {
  int32_t a, b, c;
};

// BVH::rayTriIntersect (bvh.cpp:82-111)
static bool ray_tri(const Ray& r, V3 v0, V3 v1, V3 v2, float& t, float& u, float& v) {
  V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  V3 h = cross(r.d, e2);
  float a = dot(e1, h);
  if (std::fabs(a) < 1e-8f) return false;
  float f = 1.0f / a;
  V3 s = sub(r.o, v0);
  u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  V3 q = cross(s, e1);
  v = f * dot(r.d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  t = f * dot(e2, q);
  return t > r.nearT && t < r.farT;
}

static inline float dist2_aabb(V3 p, const Node& n) {  // bvh.cpp:20-25
  float dx = std::max({n.bmin[0] - p.x, 0.0f, p.x - n.bmax[0]});
  float dy = std::max({n.bmin[1] - p.y, 0.0f, p.y - n.bmax[1]});
  float dz = std::max({n.bmin[2] - p.z, 0.0f, p.z - n.bmax[2]});
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}
static float dist2_line(V3 p, V3 a, V3 b, V3& out) {  // bvh.cpp:27-40
  V3 ab = sub(b, a), ap = sub(p, a);
  float ab2 = sq3(ab);
  if (ab2 < 1e-6f) {
    out = a;
    return sq3(ap);
  }
  float d = dot(ap, ab);
  float t = d / ab2;
  t = std::clamp(t, 0.0f, 1.0f);
  V3 proj = madd(ab, t, a);
  out = proj;
  return sq3(sub(p, proj));
}
static float dist2_tri(V3 p, V3 a, V3 b, V3 c, V3& out) {  // bvh.cpp:42-80
  V3 e1 = sub(a, b), e2 = sub(a, c);
  V3 nrm = cross(e1, e2);
  if (sq3(nrm) < 1e-6f) {
    out = a;
    return sq3(sub(p, a));
  }
  nrm = normalized(nrm);
  float dp = dot(sub(p, a), nrm);
  V3 pj = mk(fmaf(-dp, nrm.x, p.x), fmaf(-dp, nrm.y, p.y), fmaf(-dp, nrm.z, p.z));
  V3 ab = neg(e1), bc = sub(c, b), ca = e2;
  V3 c1 = cross(ab, sub(pj, a)), c2 = cross(bc, sub(pj, b)), c3 = cross(ca, sub(pj, c));
  if (dot(c1, nrm) >= 0 && dot(c2, nrm) >= 0 && dot(c3, nrm) >= 0) {
    out = pj;
    return dp * dp;
  }
  float mn = std::numeric_limits<float>::infinity(), d;
  V3 tmp;
  if ((d = dist2_line(p, a, b, tmp)) < mn) { mn = d; out = tmp; }
  if ((d = dist2_line(p, b, c, tmp)) < mn) { mn = d; out = tmp; }
  if ((d = dist2_line(p, c, a, tmp)) < mn) { mn = d; out = tmp; }
  return mn;
}

// Traversal cost of one query: nodes whose bounds were tested, triangles tested.
struct TravCost {
  uint64_t nodes = 0, tris = 0;
};

struct BVH {
  std::vector<Node> nodes;
  std::vector<uint32_t> tri;
  const V3* v = nullptr;
  const Tri* f = nullptr;
  uint32_t n_tri = 0, root = 0;

  float cen(uint32_t ti, int axis) const {
    const Tri& t = f[ti];
    return (comp(v[t.a], axis) + comp(v[t.b], axis) + comp(v[t.c], axis)) / 3.0f;
  }
  // BVH::build / buildRecursive (bvh.cpp:115-193)
  void build(const V3* verts, const Tri* faces, uint32_t n) {
    v = verts;
    f = faces;
    n_tri = n;
    nodes.clear();
    tri.clear();
    root = 0;
    if (n == 0) return;
    tri.resize(n);
    for (uint32_t i = 0; i < n; ++i) tri[i] = i;
    nodes.reserve(2 * (size_t)n);
    root = rec(0, n);
  }
  uint32_t rec(uint32_t b, uint32_t e) {
    uint32_t count = e - b;
    Node box;
    for (int a = 0; a < 3; ++a) {
      box.bmin[a] = INFINITY;
      box.bmax[a] = -INFINITY;
    }
    for (uint32_t i = b; i < e; ++i) {
      const Tri& t = f[tri[i]];
      const V3 vs[3] = {v[t.a], v[t.b], v[t.c]};
      float tmin[3] = {INFINITY, INFINITY, INFINITY}, tmax[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
          tmin[a] = std::min(tmin[a], comp(vs[k], a));
          tmax[a] = std::max(tmax[a], comp(vs[k], a));
        }
      for (int a = 0; a < 3; ++a) {
        box.bmin[a] = std::min(box.bmin[a], tmin[a]);
        box.bmax[a] = std::max(box.bmax[a], tmax[a]);
      }
    }
    if (count <= 4) {
      box.left = b;
      box.right = count | LEAF;
      nodes.push_back(box);
      return (uint32_t)nodes.size() - 1;
    }
    int axis = longest_axis(box);
    float smin = INFINITY, smax = -INFINITY;
    for (uint32_t i = b; i < e; ++i) {
      float c = cen(tri[i], axis);
      smin = std::min(smin, c);
      smax = std::max(smax, c);
    }
    float split = (smin + smax) * 0.5f;
    auto mid = std::partition(tri.begin() + b, tri.begin() + e,
                              [&](uint32_t ti) { return cen(ti, axis) < split; });
    uint32_t m = (uint32_t)(mid - tri.begin());
    if (m == b || m == e) m = b + count / 2;
    uint32_t l = rec(b, m);
    uint32_t r = rec(m, e);
    box.left = l;
    box.right = r;
    nodes.push_back(box);
    return (uint32_t)nodes.size() - 1;
  }
  void near_far(const Node& nd, const Ray& r, uint32_t& first, uint32_t& second) const {
    const Node& ln = nodes[nd.left];
    const Node& rn = nodes[nd.right];
    int axis = longest_axis(nd);
    float lc = centroid(ln, axis), rc = centroid(rn, axis);
    bool nearFirst = (comp(r.d, axis) > 0) ? (lc < rc) : (lc > rc);
    first = nearFirst ? nd.left : nd.right;
    second = nearFirst ? nd.right : nd.left;
  }
  // BVH::intersect (bvh.cpp:197-250); narrows r.farT.  cost (optional) counts
  // the nodes whose bounds are tested and the triangles tested.
  bool intersect(Ray& r, uint32_t& ti_out, float& u_out, float& v_out,
                 TravCost* cost = nullptr) const {
    if (nodes.empty()) return false;
    bool hit = false;
    uint32_t st[64];
    int sp = 0;
    st[sp++] = root;
    while (sp > 0) {
      const Node& nd = nodes[st[--sp]];
      if (cost) ++cost->nodes;
      if (!aabb_hit(nd, r)) continue;
      if (nd.right & LEAF) {
        uint32_t s = nd.left, c = nd.right & ~LEAF;
        for (uint32_t i = s; i < s + c; ++i) {
          if (cost) ++cost->tris;
          uint32_t ti = tri[i];
          float t, u, vv;
          if (ray_tri(r, v[f[ti].a], v[f[ti].b], v[f[ti].c], t, u, vv)) {
            r.farT = t;
            ti_out = ti;
            u_out = u;
            v_out = vv;
            hit = true;
          }
        }
      } else {
        uint32_t a, b;
        near_far(nd, r, a, b);
        if (sp <= 62) {  // 64-entry stack
          st[sp++] = b;
          st[sp++] = a;
        }
      }
    }
    return hit;
  }
  // BVH::occluded (bvh.cpp:254-300): starts at node 0, not the root
  bool occluded(const Ray& r) const {
    if (nodes.empty()) return false;
    uint32_t st[64];
    int sp = 0;
    st[sp++] = 0;
    while (sp > 0) {
      const Node& nd = nodes[st[--sp]];
      if (!aabb_hit(nd, r)) continue;
      if (nd.right & LEAF) {
        uint32_t s = nd.left, c = nd.right & ~LEAF;
        for (uint32_t i = s; i < s + c; ++i) {
          uint32_t ti = tri[i];
          float t, u, vv;
          if (ray_tri(r, v[f[ti].a], v[f[ti].b], v[f[ti].c], t, u, vv)) return true;
        }
      } else {
        uint32_t a, b;
        near_far(nd, r, a, b);
        if (sp <= 62) {  // 64-entry stack
          st[sp++] = b;
          st[sp++] = a;
        }
      }
    }
    return false;
  }
  // BVH::closest (bvh.cpp:302-353); cost as for intersect, where expanding a
  // node tests the bounds of both children
  float closest(V3 p, uint32_t& ti_out, V3& out, TravCost* cost = nullptr) const {
    if (nodes.empty()) return 0.0f;
    uint32_t st[64];
    int sp = 0;
    st[sp++] = root;
    float md = std::numeric_limits<float>::max();
    while (sp > 0) {
      const Node& nd = nodes[st[--sp]];
      if (cost) ++cost->nodes;
      if (dist2_aabb(p, nd) >= md) continue;
      if (nd.right & LEAF) {
        uint32_t s = nd.left, c = nd.right & ~LEAF;
        for (uint32_t i = s; i < s + c; ++i) {
          if (cost) ++cost->tris;
          uint32_t ti = tri[i];
          V3 tmp{0, 0, 0};
          float d = dist2_tri(p, v[f[ti].a], v[f[ti].b], v[f[ti].c], tmp);
          if (d < md) {
            md = d;
            ti_out = ti;
            out = tmp;
          }
        }
      } else {
        float dl = dist2_aabb(p, nodes[nd.left]);
        float dr = dist2_aabb(p, nodes[nd.right]);
        if (cost) cost->nodes += 2;
        if (sp > 62) continue;
        if (dl < dr) {
          if (dr < md) st[sp++] = nd.right;
          if (dl < md) st[sp++] = nd.left;
        } else {
          if (dl < md) st[sp++] = nd.left;
          if (dr < md) st[sp++] = nd.right;
        }
      }
    }
    return sqrtf(md);
  }
};

// ---------------------------------------------------------------- texture (texture.cpp:52-69)
struct Tex {
  int w = 0, h = 0;
  std::vector<float> rgb;
  bool present() const { return w > 0 && h > 0; }
  V3 at(float u, float v, int* texel = nullptr) const {
    u = u - std::floor(u);
    v = v - std::floor(v);
    v = 1.0f - v;
    int x = (int)(u * w) % w;
    int y = (int)(v * h) % h;
    if (x < 0) x += w;
    if (y < 0) y += h;
    int i = y * w + x;
    if (texel) *texel = i;
    return mk(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
  }
};
static Tex make_tex(const mupsi_texture_desc& d) {
  Tex t;
  if (!d.rgb || d.width <= 0 || d.height <= 0) return t;
  t.w = d.width;
  t.h = d.height;
  t.rgb.assign(d.rgb, d.rgb + (size_t)d.width * d.height * 3);
  return t;
}

// ---------------------------------------------------------------- mesh (mesh.cpp)
struct Mesh {
  std::vector<V3> verts, normals, fnormals;
  std::vector<float> tex;  // texcoords x 2
  std::vector<Tri> faces, ftex;
  std::vector<float> cdf;
  float total_area = 0;
  BVH bvh;
  Xf xf;

  void init(const mupsi_mesh_desc& d) {
    verts.resize(d.n_vertices);
    for (int i = 0; i < d.n_vertices; ++i)
      verts[i] = mk(d.vertices[3 * i], d.vertices[3 * i + 1], d.vertices[3 * i + 2]);
    if (d.normals && d.n_normals > 0) {
      normals.resize(d.n_normals);
      for (int i = 0; i < d.n_normals; ++i)
        normals[i] = mk(d.normals[3 * i], d.normals[3 * i + 1], d.normals[3 * i + 2]);
    }
    if (d.texcoords && d.n_texcoords > 0) tex.assign(d.texcoords, d.texcoords + 2 * d.n_texcoords);
    faces.resize(d.n_faces);
    for (int i = 0; i < d.n_faces; ++i)
      faces[i] = Tri{d.faces[3 * i], d.faces[3 * i + 1], d.faces[3 * i + 2]};
    if (d.face_tex) {
      ftex.resize(d.n_faces);
      for (int i = 0; i < d.n_faces; ++i)
        ftex[i] = Tri{d.face_tex[3 * i], d.face_tex[3 * i + 1], d.face_tex[3 * i + 2]};
    }
    xf = make_xf(mk(d.position[0], d.position[1], d.position[2]), d.scale);
    // prepareForRender (mesh.cpp:299-319): face normals, BVH, area CDF
    fnormals.resize(faces.size());
    for (size_t i = 0; i < faces.size(); ++i) {
      V3 e1 = sub(verts[faces[i].b], verts[faces[i].a]);
      V3 e2 = sub(verts[faces[i].c], verts[faces[i].a]);
      V3 n = cross(e1, e2);
      float len = norm(n);
      fnormals[i] = len > 1e-12f ? divs(n, len) : mk(0, 0, 0);
    }
    bvh.build(verts.data(), faces.data(), (uint32_t)faces.size());
    total_area = 0;
    cdf.resize(faces.size());
    for (size_t i = 0; i < faces.size(); ++i) {
      V3 e1 = sub(verts[faces[i].b], verts[faces[i].a]);
      V3 e2 = sub(verts[faces[i].c], verts[faces[i].a]);
      total_area += 0.5f * norm(cross(e1, e2));
      cdf[i] = total_area;
    }
    if (total_area > 0)
      for (auto& c : cdf) c /= total_area;
  }
  V3 interp_normal(uint32_t fi, float u, float v) const {  // mesh.cpp:163-172
    if (normals.empty()) return fnormals[fi];
    const Tri& f = faces[fi];
    float w = 1.0f - u - v;
    V3 n0 = normals[f.a], n1 = normals[f.b], n2 = normals[f.c];
    V3 n = mk(fmaf(v, n2.x, fmaf(u, n1.x, w * n0.x)), fmaf(v, n2.y, fmaf(u, n1.y, w * n0.y)),
              fmaf(v, n2.z, fmaf(u, n1.z, w * n0.z)));
    float len = norm(n);
    return len > 1e-12f ? divs(n, len) : fnormals[fi];
  }
  void interp_uv(uint32_t fi, float u, float v, float& ou, float& ov) const {  // mesh.cpp:174-181
    if (tex.empty() || ftex.empty()) {
      ou = u;
      ov = v;
      return;
    }
    const Tri& t = ftex[fi];
    float w = 1.0f - u - v;
    ou = fmaf(v, tex[2 * t.c], fmaf(u, tex[2 * t.b], w * tex[2 * t.a]));
    ov = fmaf(v, tex[2 * t.c + 1], fmaf(u, tex[2 * t.b + 1], w * tex[2 * t.a + 1]));
  }
  // Mesh::intersect (mesh.cpp:185-209): narrows ray.farT on hit
  bool intersect(Ray& ray, uint32_t& ti, float& u, float& v) const {
    Ray lr = ray;
    float tcorr = 1.0f;
    if (!xf.identity) {
      V3 lo = xf_to_local(xf, ray.o);
      V3 ld = xf_dir_to_local(xf, ray.d);
      float len = norm(ld);
      tcorr = norm(xf_dir_to_world(xf, divs(ld, len)));
      lr = make_ray(lo, ld);
      lr.nearT = ray.nearT * len;
      lr.farT = ray.farT * len;
    }
    bool hit = bvh.intersect(lr, ti, u, v);
    if (hit) ray.farT = lr.farT * tcorr;
    return hit;
  }
  bool occluded(const Ray& ray) const {  // mesh.cpp:236-250
    if (xf.identity) return bvh.occluded(ray);
    V3 lo = xf_to_local(xf, ray.o);
    V3 ld = xf_dir_to_local(xf, ray.d);
    float len = norm(ld);
    Ray lr = make_ray(lo, ld);
    lr.nearT = ray.nearT * len;
    lr.farT = ray.farT * len;
    return bvh.occluded(lr);
  }
  // Mesh::distToPoint (mesh.cpp:325-339)
  float dist_to_point(V3 p, uint32_t& ti, V3& out, TravCost* cost = nullptr) const {
    V3 pl = xf_to_local(xf, p);
    V3 ol = mk(0, 0, 0);
    ti = 0;
    bvh.closest(pl, ti, ol, cost);
    out = xf_to_world(xf, ol);
    return norm(sub(out, p));
  }
  // Mesh::inside (mesh.cpp:341-362)
  bool inside(V3 p, TravCost* cost = nullptr) const {
    V3 pl = xf_to_local(xf, p);
    Ray r = make_ray(pl, mk(1, 0, 0));
    int hits = 0;
    uint32_t ti;
    float u, v;
    if (bvh.intersect(r, ti, u, v, cost)) {
      ++hits;
      while (r.farT < std::numeric_limits<float>::infinity()) {
        r.nearT = r.farT + 1e-4f;
        r.farT = std::numeric_limits<float>::infinity();
        if (!bvh.intersect(r, ti, u, v, cost)) break;
        ++hits;
      }
    }
    return (hits % 2) == 1;
  }
  V3 world_normal(V3 n) const {  // normalTransform_ * n, normalized
    if (xf.identity) return n;
    return normalized(mul(n, xf.inv_s));
  }
};

// ---------------------------------------------------------------- scene
struct Prim {
  int type, bsdf;
  Xf xf;  // sphere placement
  bool has_em;
  V3 em;
  bool int_medium;
  int mesh = -1;
  Tex albedo;
};

struct Cam {
  V3 pos, dir, up, right, tup;
  float fov, aspect;
  int w, h;
};

struct Cond {
  bool active = false;
  V3 C{0, 0, 0};
  float u = 0;
  V3 g{0, 0, 0};
};

// ExternalGPMedium grids (GridMeanFunction, meanfunction.cpp:41-86)
struct Grid {
  int res = 0;
  V3 bmin{0, 0, 0}, inv{0, 0, 0};
  float noise_scale = 1.0f;
  std::vector<float> mean, var;
};

struct Scene {
  Cam cam;
  int max_bounce = 3, gp_mode = 0, skip_medium = 0;
  bool has_gp = false;       // a medium of medium_kind exists
  int medium_kind = MUPSI_MEDIUM_GP;
  Grid grid;
  int mean_type = 0;
  V3 mean_c{0, 0, 0};
  float mean_r = 70;
  Mesh mean_mesh;
  GP gp;
  Tex sky;
  std::vector<Prim> prims;
  std::vector<Mesh> meshes;
  float eps = 1e-4f;
  float near_eps = (float)((1e-4) * 15);
};

// Camera (camera.h:14-19, camera.cpp:5-19)
static Cam make_cam(const mupsi_scene_desc& d) {
  Cam c;
  c.pos = mk(d.cam_pos[0], d.cam_pos[1], d.cam_pos[2]);
  c.dir = normalized(mk(d.cam_dir[0], d.cam_dir[1], d.cam_dir[2]));
  c.up = normalized(mk(d.cam_up[0], d.cam_up[1], d.cam_up[2]));
  c.fov = d.fov;
  c.w = d.width;
  c.h = d.height;
  c.aspect = (float)d.width / (float)d.height;
  c.right = normalized(cross(c.dir, c.up));
  c.tup = normalized(cross(c.right, c.dir));
  return c;
}
static V3 cam_dir(const Cam& c, int x, int y) {
  float cx = (x + 0.5f) / c.w;
  float cy = (y + 0.5f) / c.h;
  float tf = tanf(c.fov / 2.0f);
  float nx = fmaf(2.0f, cx, -1.0f) * c.aspect * tf;
  float ny = fmaf(2.0f, cy, -1.0f) * tf;
  V3 rc = normalized(mk(nx, ny, 1.0f));
  V3 w = mk(fmaf(c.dir.x, rc.z, fmaf(c.tup.x, rc.y, c.right.x * rc.x)),
            fmaf(c.dir.y, rc.z, fmaf(c.tup.y, rc.y, c.right.y * rc.x)),
            fmaf(c.dir.z, rc.z, fmaf(c.tup.z, rc.y, c.right.z * rc.x)));
  return normalized(normalized(w));
}

// Mean functions (meanfunction.cpp:13-39)
static float mean_eval(const Scene& s, V3 p) {
  if (s.mean_type == MUPSI_MEAN_SPHERE) return norm(sub(p, s.mean_c)) - s.mean_r;
  float sg = s.mean_mesh.inside(p) ? -1.0f : 1.0f;
  uint32_t ti;
  V3 o;
  return sg * s.mean_mesh.dist_to_point(p, ti, o);
}
static V3 mean_grad(const Scene& s, V3 p) {
  if (s.mean_type == MUPSI_MEAN_SPHERE) return normalized(sub(p, s.mean_c));
  float sg = s.mean_mesh.inside(p) ? -1.0f : 1.0f;
  uint32_t ti;
  V3 o;
  float d = s.mean_mesh.dist_to_point(p, ti, o);
  if (d < 1e-6f) return s.mean_mesh.world_normal(s.mean_mesh.fnormals[ti]);
  return normalized(mul(sub(p, o), sg));
}
static double mean_eval_d(const Scene& s, V3 p) {
  if (s.mean_type == MUPSI_MEAN_SPHERE) {
    double dx = (double)p.x - s.mean_c.x, dy = (double)p.y - s.mean_c.y,
           dz = (double)p.z - s.mean_c.z;
    return std::sqrt(dx * dx + dy * dy + dz * dz) - (double)s.mean_r;
  }
  return mean_eval(s, p);
}
static D3 mean_grad_d(const Scene& s, V3 p) {
  if (s.mean_type == MUPSI_MEAN_SPHERE) {
    double dx = (double)p.x - s.mean_c.x, dy = (double)p.y - s.mean_c.y,
           dz = (double)p.z - s.mean_c.z;
    double l = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (l > 0) return D3{dx / l, dy / l, dz / l};
    return D3{dx, dy, dz};
  }
  V3 g = mean_grad(s, p);
  return D3{g.x, g.y, g.z};
}

// GPMedium::evalWithGradient(p, seed, state) (gpmedium.cpp:19-41), FP32
static F4 field_f32(const Scene& s, V3 p, uint32_t seed, const Cond& c, float* l1 = nullptr) {
  float mu = mean_eval(s, p);
  V3 mg = mean_grad(s, p);
  F4 ps = raw_noise_f32(s.gp, p, seed, l1);
  F4 f{mu + ps.v, mg.x + ps.gx, mg.y + ps.gy, mg.z + ps.gz};
  if (c.active) {
    V3 hg = kern_hgrad(s.gp, c.C, p);
    f.v = f.v + fmaf(c.u, kern_h(s.gp, c.C, p), dot(c.g, hg));
    V3 Hg = kern_hess_vec(s.gp, c.C, p, c.g);
    f.gx = f.gx + fmaf(c.u, hg.x, Hg.x);
    f.gy = f.gy + fmaf(c.u, hg.y, Hg.y);
    f.gz = f.gz + fmaf(c.u, hg.z, Hg.z);
  }
  return f;
}
// FP64 counterpart with the same integer decisions
static F4d field_f64(const Scene& s, V3 p, uint32_t seed, const Cond& c, double* l1 = nullptr,
                     double* mu_out = nullptr) {
  double mu = mean_eval_d(s, p);
  D3 mg = mean_grad_d(s, p);
  F4d ps = raw_noise_f64(s.gp, p, seed, l1);
  F4d f{mu + ps.v, mg.x + ps.gx, mg.y + ps.gy, mg.z + ps.gz};
  if (mu_out) *mu_out = mu;
  if (c.active) {
    double L2 = s.gp.L2;
    double rx = (double)p.x - c.C.x, ry = (double)p.y - c.C.y, rz = (double)p.z - c.C.z;
    double h = std::exp(-(rx * rx + ry * ry + rz * rz) / L2);
    double k = h * (-2.0 / L2);
    double hx = k * rx, hy = k * ry, hz = k * rz;
    f.v += c.u * h + (c.g.x * hx + c.g.y * hy + c.g.z * hz);
    double rv = rx * c.g.x + ry * c.g.y + rz * c.g.z;
    double k4 = 4.0 / (L2 * L2) * rv;
    f.gx += c.u * hx + h * (-2.0 / L2 * c.g.x + k4 * rx);
    f.gy += c.u * hy + h * (-2.0 / L2 * c.g.y + k4 * ry);
    f.gz += c.u * hz + h * (-2.0 / L2 * c.g.z + k4 * rz);
  }
  return f;
}

static inline bool sgn(float t) { return t > 0.0f; }

struct MarchOut {
  bool hit = false;
  float t = 0;
  V3 p{0, 0, 0}, n{0, 0, 0};
  int steps = 0;
  double amb = INFINITY, cell = INFINITY;
};

static double cell_margin(const GP& g, V3 p) {
  const float q[3] = {p.x / g.kr, p.y / g.kr, p.z / g.kr};
  double m = INFINITY;
  for (int a = 0; a < 3; ++a) {
    double f = (double)q[a] - std::trunc((double)q[a]);
    m = std::min(m, std::min(std::fabs(f), 1.0 - std::fabs(f)));
  }
  return m;
}

// GPMedium::sampleDistance (mode 0) / transmittance (mode 1)  (gpmedium.cpp:46-129).
// The FP32 path also runs the reference's discarded lerp/0.9 refinement so the
// CPU baseline does the same work as the reference.
static MarchOut sdf_march(const Scene& s, const Ray& ray);
static MarchOut grid_march(const Scene& s, const Ray& ray, uint32_t seed);
static MarchOut march(const Scene& s, int mode, const Ray& ray, uint32_t seed, const Cond& c,
                      bool fp64, bool track) {
  if (s.medium_kind == MUPSI_MEDIUM_SDF) return sdf_march(s, ray);
  if (s.medium_kind == MUPSI_MEDIUM_GRID) return grid_march(s, ray, seed);
  MarchOut m;
  float nt = ray.nearT, ft = std::min(ray.farT, 2000.0f);
  float dt = std::min(0.5f, (ft - nt) / 100);
  float t = nt;
  double fprev_d = 0;
  float fprev;
  if (fp64) {
    fprev_d = field_f64(s, ray.o, seed, c).v;
    fprev = (float)fprev_d;
  } else {
    fprev = field_f32(s, ray.o, seed, c).v;
  }
  bool sp = fp64 ? (fprev_d > 0) : sgn(fprev);
  while (t < ft) {
    V3 p = madd(ray.d, t, ray.o);
    ++m.steps;
    bool differ;
    double fc_d = 0;
    float fc = 0;
    if (fp64) {
      double l1 = 0, mu = 0;
      fc_d = field_f64(s, p, seed, c, &l1, &mu).v;
      differ = ((fc_d > 0) != sp) && std::fabs(fprev_d - fc_d) > 1e-6;
      if (track) {
        m.amb = std::min(m.amb, std::fabs(fc_d) / (1.0 + std::fabs(mu) + l1));
        m.cell = std::min(m.cell, cell_margin(s.gp, p));
      }
    } else {
      fc = field_f32(s, p, seed, c).v;
      differ = sgn(fc) != sp && (double)std::fabs(fprev - fc) > 1e-6;
      if (differ && mode == 0) {  // discarded refinement (gpmedium.cpp:66-79)
        float factor = fprev / (fprev - fc);
        do {
          float tt = std::lerp(t - dt, t, factor);
          float ftst = field_f32(s, madd(ray.d, tt, ray.o), seed, c).v;
          if (sgn(ftst) == sp) break;
          factor = (float)(factor * 0.9);
        } while (factor >= 0.01);
      }
    }
    if (differ) {
      m.hit = true;
      m.t = t;
      if (mode == 0) {
        m.p = madd(ray.d, t, ray.o);
        if (fp64) {
          F4d f = field_f64(s, m.p, seed, c);
          double l = std::sqrt(f.gx * f.gx + f.gy * f.gy + f.gz * f.gz);
          m.n = mk((float)(f.gx / l), (float)(f.gy / l), (float)(f.gz / l));
        } else {
          F4 f = field_f32(s, m.p, seed, c);
          m.n = normalized(mk(f.gx, f.gy, f.gz));
        }
      }
      return m;
    }
    if (!(t + dt > t)) return m;  // the t grid cannot advance (the reference would not terminate)
    t += dt;
  }
  return m;
}

// GPMedium::sampleCondition (gpmedium.cpp:133-162)
static Cond condition_f32(const Scene& s, V3 C, V3 n, uint32_t seed_next) {
  Cond c;
  if (s.gp_mode != MUPSI_GP_RENEWAL_PLUS) return c;
  c.active = true;
  c.C = C;
  F4 ps = raw_noise_f32(s.gp, C, seed_next, nullptr);
  c.u = -mean_eval(s, C) / s.gp.sigma - ps.v;
  V3 gm = mean_grad(s, C);
  float k = -s.gp.L2 / 2.0f;
  c.g = mk(k * ((n.x - gm.x) - ps.gx), k * ((n.y - gm.y) - ps.gy), k * ((n.z - gm.z) - ps.gz));
  return c;
}
static void condition_f64(const Scene& s, V3 C, V3 n, uint32_t seed_next, double* u, double* g) {
  F4d ps = raw_noise_f64(s.gp, C, seed_next, nullptr);
  double mu = mean_eval_d(s, C);
  D3 gm = mean_grad_d(s, C);
  double k = -(double)s.gp.L2 / 2.0;
  *u = -mu / (double)s.gp.sigma - ps.v;
  g[0] = k * ((double)n.x - gm.x - ps.gx);
  g[1] = k * ((double)n.y - gm.y - ps.gy);
  g[2] = k * ((double)n.z - gm.z - ps.gz);
}

// ---------------------------------------------------------------- SDF / grid media
// GridMeanFunction::eval (meanfunction.cpp:56-76): frac = (p - min) * inv - 0.5,
// indices clamped to [0, res - 2], weights clamped to [0, 1], data x-major;
// every a*b + c*d as fmaf(c, d, a*b).
static float grid_eval(const Grid& g, const std::vector<float>& data, V3 p) {
  float fx = fmaf(p.x - g.bmin.x, g.inv.x, -0.5f);
  float fy = fmaf(p.y - g.bmin.y, g.inv.y, -0.5f);
  float fz = fmaf(p.z - g.bmin.z, g.inv.z, -0.5f);
  int i0 = std::max(0, std::min(g.res - 2, (int)std::floor(fx)));
  int j0 = std::max(0, std::min(g.res - 2, (int)std::floor(fy)));
  int k0 = std::max(0, std::min(g.res - 2, (int)std::floor(fz)));
  float wx = std::clamp(fx - (float)i0, 0.0f, 1.0f);
  float wy = std::clamp(fy - (float)j0, 0.0f, 1.0f);
  float wz = std::clamp(fz - (float)k0, 0.0f, 1.0f);
  auto at = [&](int i, int j, int k) {
    return data[(size_t)i * g.res * g.res + (size_t)j * g.res + k];
  };
  float c00 = fmaf(at(i0 + 1, j0, k0), wx, at(i0, j0, k0) * (1 - wx));
  float c01 = fmaf(at(i0 + 1, j0, k0 + 1), wx, at(i0, j0, k0 + 1) * (1 - wx));
  float c10 = fmaf(at(i0 + 1, j0 + 1, k0), wx, at(i0, j0 + 1, k0) * (1 - wx));
  float c11 = fmaf(at(i0 + 1, j0 + 1, k0 + 1), wx, at(i0, j0 + 1, k0 + 1) * (1 - wx));
  float c0 = fmaf(c10, wy, c00 * (1 - wy));
  float c1 = fmaf(c11, wy, c01 * (1 - wy));
  return fmaf(c1, wz, c0 * (1 - wz));
}
// GridMeanFunction::gradient (meanfunction.cpp:78-84)
static V3 grid_grad(const Grid& g, const std::vector<float>& data, V3 p) {
  const float e = 1e-4f;
  float dx = grid_eval(g, data, mk(p.x + e, p.y, p.z)) - grid_eval(g, data, mk(p.x - e, p.y, p.z));
  float dy = grid_eval(g, data, mk(p.x, p.y + e, p.z)) - grid_eval(g, data, mk(p.x, p.y - e, p.z));
  float dz = grid_eval(g, data, mk(p.x, p.y, p.z + e)) - grid_eval(g, data, mk(p.x, p.y, p.z - e));
  return divs(mk(dx, dy, dz), 2.0f * e);
}
// hash11 (externalgpmedium.cpp:8-17)
static float hash11(float x, float y, float z, uint32_t seed) {
  uint32_t h = seed, b;
  std::memcpy(&b, &x, 4);
  h = h * 0x9e3779b1u + b;
  std::memcpy(&b, &y, 4);
  h = h * 0x9e3779b1u + b;
  std::memcpy(&b, &z, 4);
  h = h * 0x9e3779b1u + b;
  h = (h ^ (h >> 16)) * 0x85ebca6bu;
  h = (h ^ (h >> 13)) * 0xc2b2ae35u;
  h = h ^ (h >> 16);
  return (float)(h & 0xFFFFFFu) / (float)0xFFFFFFu * 2.0f - 1.0f;
}
// ExternalGPMedium::evalWithGradient (externalgpmedium.cpp:31-58)
static F4 grid_medium(const Scene& s, V3 p, uint32_t seed) {
  const Grid& g = s.grid;
  float mu = grid_eval(g, g.mean, p);
  V3 mg = grid_grad(g, g.mean, p);
  float noise = 0.0f;
  V3 ng{0, 0, 0};
  if (!g.var.empty() && g.noise_scale > 0.0f) {
    float var = grid_eval(g, g.var, p);
    if (var > 0.0f) {
      float amp = g.noise_scale * std::sqrt(var);
      noise = amp * hash11(p.x, p.y, p.z, seed);
      float e = 0.01f;
      ng.x = amp * (hash11(p.x + e, p.y, p.z, seed) - hash11(p.x - e, p.y, p.z, seed)) / (2.0f * e);
      ng.y = amp * (hash11(p.x, p.y + e, p.z, seed) - hash11(p.x, p.y - e, p.z, seed)) / (2.0f * e);
      ng.z = amp * (hash11(p.x, p.y, p.z + e, seed) - hash11(p.x, p.y, p.z - e, seed)) / (2.0f * e);
    }
  }
  return F4{mu + noise, mg.x + ng.x, mg.y + ng.y, mg.z + ng.z};
}
static float std_max(float a, float b) { return std::max(a, b); }
// SDFMedium::sampleDistance / transmittance (sdfmedium.cpp:16-98)
static MarchOut sdf_march(const Scene& s, const Ray& ray) {
  MarchOut m;
  float nt = ray.nearT, ft = std::min(ray.farT, 2000.0f);
  float t = nt;
  float fc = mean_eval(s, ray.o);
  int iter = 0;
  while (t < ft && iter < 512) {
    V3 p = madd(ray.d, t, ray.o);
    float fp = fc, tp = t;
    fc = mean_eval(s, p);
    if (fp * fc < 0.0f) {
      float lo = tp, hi = t, flo = fp;
      for (int b = 0; b < 16; b++) {
        float mid = (lo + hi) * 0.5f;
        float fm = mean_eval(s, madd(ray.d, mid, ray.o));
        if (flo * fm <= 0.0f) {
          hi = mid;
        } else {
          lo = mid;
          flo = fm;
        }
      }
      m.hit = true;
      m.t = hi;
      m.p = madd(ray.d, hi, ray.o);
      m.n = normalized(mean_grad(s, m.p));
      m.steps = iter + 1;
      return m;
    }
    t += std_max(std::fabs(fc), 1e-3f);
    iter++;
  }
  m.steps = iter;
  return m;
}
// ExternalGPMedium::sampleDistance / transmittance (externalgpmedium.cpp:68-130)
static MarchOut grid_march(const Scene& s, const Ray& ray, uint32_t seed) {
  MarchOut m;
  float nt = ray.nearT, ft = std::min(ray.farT, 2000.0f);
  const float min_step = 0.005f;
  float t = nt;
  float f_prev = grid_medium(s, ray.o, seed).v;
  int iter = 0;
  for (; iter < 1024 && t < ft; ++iter) {
    V3 p = madd(ray.d, t, ray.o);
    float f_c = grid_medium(s, p, seed).v;
    if (sgn(f_c) != sgn(f_prev) && std::fabs(f_prev - f_c) > 1e-6f) {
      float lo = t - std_max(std::fabs(f_prev), min_step);
      float hi = t;
      for (int b = 0; b < 16; ++b) {
        float mid = (lo + hi) * 0.5f;
        float f_mid = grid_medium(s, madd(ray.d, mid, ray.o), seed).v;
        if (sgn(f_mid) == sgn(f_prev)) lo = mid;
        else hi = mid;
      }
      m.hit = true;
      m.t = hi;
      m.p = madd(ray.d, hi, ray.o);
      F4 f = grid_medium(s, m.p, seed);
      m.n = normalized(mk(f.gx, f.gy, f.gz));
      m.steps = iter + 1;
      return m;
    }
    f_prev = f_c;
    t += std_max(std::fabs(f_c), min_step);
  }
  m.steps = iter;
  return m;
}

// ---------------------------------------------------------------- primitives
// intersectUnitSphere (sphere.cpp:11-30)
static float unit_sphere_t(float nearT, float farT, V3 O, V3 D) {
  float a = sq3(D);
  float b = dot(O, D);
  float c = sq3(O) - 1.0f;
  float disc = fmaf(b, b, -(a * c));
  if (disc < 0.0f) return -1.0f;
  float sd = sqrtf(disc);
  float t = (-b - sd) / a;
  if (t > nearT && t < farT) return t;
  t = (-b + sd) / a;
  if (t > nearT && t < farT) return t;
  return -1.0f;
}
static void sphere_local(const Prim& pr, const Ray& r, V3& O, V3& D) {
  O = r.o;
  D = r.d;
  if (!pr.xf.identity) {
    O = xf_to_local(pr.xf, r.o);
    D = xf_dir_to_local(pr.xf, r.d);
  }
}

struct Hit {
  int prim = -1;
  uint32_t tri = 0;
  float u = 0, v = 0;
};

// Scene::intersect (scene.cpp:25-39): returns primitive index or -1
static int scene_intersect(const Scene& s, Ray& ray, V3& P, V3& Ng, float& uu, float& vv) {
  float mint = std::numeric_limits<float>::max();
  int hitp = -1;
  uint32_t hit_tri = 0;
  float hit_u = 0, hit_v = 0;
  for (size_t i = 0; i < s.prims.size(); ++i) {
    const Prim& pr = s.prims[i];
    bool h = false;
    uint32_t ti = 0;
    float u = 0, v = 0;
    if (pr.type == MUPSI_PRIM_SPHERE) {
      V3 O, D;
      sphere_local(pr, ray, O, D);
      float t = unit_sphere_t(ray.nearT, ray.farT, O, D);
      if (t >= 0.0f) {
        ray.farT = t;
        h = true;
      }
    } else {
      h = s.meshes[pr.mesh].intersect(ray, ti, u, v);
    }
    if ((h || hitp >= 0) && ray.farT < mint) {
      if (h) {
        mint = ray.farT;
        hitp = (int)i;
        hit_tri = ti;
        hit_u = u;
        hit_v = v;
      }
    }
  }
  if (hitp < 0) return -1;
  P = madd(ray.d, ray.farT, ray.o);
  const Prim& pr = s.prims[hitp];
  if (pr.type == MUPSI_PRIM_SPHERE) {  // Sphere::intersectInfo (sphere.cpp:35-54)
    V3 pl = pr.xf.identity ? P : xf_to_local(pr.xf, P);
    V3 ln = normalized(pl);
    Ng = pr.xf.identity ? ln : normalized(mul(ln, pr.xf.inv_s));
    uu = 0;
    vv = 0;  // sphere uv only feeds constant textures
  } else {  // Mesh::intersectInfo (mesh.cpp:211-234)
    const Mesh& m = s.meshes[pr.mesh];
    Ng = m.world_normal(m.interp_normal(hit_tri, hit_u, hit_v));
    m.interp_uv(hit_tri, hit_u, hit_v, uu, vv);
  }
  return hitp;
}

// Scene::occluded (scene.cpp:41-48)
static bool scene_occluded(const Scene& s, const Ray& r) {
  for (const Prim& pr : s.prims) {
    if (pr.type == MUPSI_PRIM_SPHERE) {
      V3 O, D;
      sphere_local(pr, r, O, D);
      if (unit_sphere_t(r.nearT, r.farT, O, D) >= 0.0f) return true;
    } else if (s.meshes[pr.mesh].occluded(r)) {
      return true;
    }
  }
  return false;
}

struct LightSample {
  V3 w, d;
  float dist, pdf;
};

// TangentFrame(n) (sample.h:25-35) and toGlobal (sample.h:45-47)
struct Frame {
  V3 n, t, b;
};
static Frame frame(V3 nin) {
  Frame f;
  f.n = normalized(nin);
  float sign = copysignf(1.0f, f.n.z);
  const float a = -1.0f / (sign + f.n.z);
  const float b = f.n.x * f.n.y * a;
  f.t = normalized(mk(fmaf(sign * f.n.x * f.n.x, a, 1.0f), sign * b, -sign * f.n.x));
  f.b = normalized(mk(b, fmaf(f.n.y * f.n.y, a, sign), -f.n.y));
  return f;
}
static V3 to_global(const Frame& f, V3 p) {
  return mk(fmaf(f.n.x, p.z, fmaf(f.b.x, p.y, f.t.x * p.x)),
            fmaf(f.n.y, p.z, fmaf(f.b.y, p.y, f.t.y * p.x)),
            fmaf(f.n.z, p.z, fmaf(f.b.z, p.y, f.t.z * p.x)));
}

// Sphere::sampleDirect (sphere.cpp:91-133)
static bool sphere_sample_direct(const Prim& pr, V3 pw, Pcg& rng, LightSample& ls) {
  V3 pl = pr.xf.identity ? pw : xf_to_local(pr.xf, pw);
  V3 dir = neg(pl);
  float dist2 = sq3(dir);
  float d2 = dist2 - 1.0f;
  if (d2 < 1e-6f) return false;
  float d = sqrtf(d2);
  float dist = sqrtf(dist2);
  dir = divs(dir, dist);
  float cosT = d / dist;
  float x1 = rng.next1D();
  float x2 = rng.next1D();
  float phi = (float)((double)(x1 * 2) * M_PI);
  float z = fmaf(x2, 1.0f - cosT, cosT);
  float r = sqrtf(std::max(fmaf(-z, z, 1.0f), 0.0f));
  V3 sd = mk(cosf(phi) * r, sinf(phi) * r, z);
  float B = dist * sd.z;
  float det = sqrtf(std::max(fmaf(B, B, -d2), 0.0f));
  ls.dist = B - det;
  ls.d = to_global(frame(dir), sd);
  if (!pr.xf.identity) {
    V3 spl = madd(ls.d, ls.dist, pl);
    V3 spw = xf_to_world(pr.xf, spl);
    V3 wd = sub(spw, pw);
    ls.dist = norm(wd);
    ls.d = divs(wd, ls.dist);
  }
  ls.pdf = (float)(1.0f / (2.0f * M_PI * (1.0f - cosT)));
  V3 em = pr.has_em ? pr.em : mk(0, 0, 0);
  ls.w = divs(em, std::max(ls.pdf, 1e-6f));
  return true;
}

// Mesh::sampleDirect (mesh.cpp:252-296): std::lower_bound on the normalised area
// CDF picks the face, (1 - sqrt(r1), r2 sqrt(r1)) the barycentrics; the pdf is
// the area measure 1/totalArea (local space) with no cos/dist^2 conversion.
static bool mesh_sample_direct(const Prim& pr, const Mesh& m, V3 p, Pcg& rng, LightSample& ls) {
  if (m.cdf.empty() || !(m.total_area > 0.0f) || m.faces.empty()) return false;
  const float r = rng.next1D();
  size_t fi = (size_t)(std::lower_bound(m.cdf.begin(), m.cdf.end(), r) - m.cdf.begin());
  if (fi >= m.faces.size()) fi = m.faces.size() - 1;
  const float r1 = rng.next1D();
  const float r2 = rng.next1D();
  const float sr1 = sqrtf(r1);
  const float bu = 1.0f - sr1, bv = r2 * sr1;
  const Tri& f = m.faces[fi];
  const V3 v0 = m.verts[f.a];
  const V3 e1 = sub(m.verts[f.b], v0), e2 = sub(m.verts[f.c], v0);
  const V3 pl = mk(fmaf(e2.x, bv, fmaf(e1.x, bu, v0.x)), fmaf(e2.y, bv, fmaf(e1.y, bu, v0.y)),
                   fmaf(e2.z, bv, fmaf(e1.z, bu, v0.z)));
  const V3 pw = m.xf.identity ? pl : xf_to_world(m.xf, pl);
  const V3 dir = sub(pw, p);
  const float dist = sqrtf(sq3(dir));
  ls.d = divs(dir, dist);
  ls.dist = dist;
  ls.pdf = 1.0f / m.total_area;
  const V3 em = pr.has_em ? pr.em : mk(0, 0, 0);
  ls.w = divs(em, std::max(ls.pdf, 1e-6f));
  return true;
}

// Scene::chooseLight (scene.cpp:50-61); *chosen (if given) receives the
// primitive index of the light drawn, -1 when none is.
static bool choose_light(const Scene& s, V3 p, Pcg& rng, LightSample& ls, int* chosen = nullptr) {
  if (chosen) *chosen = -1;
  if (s.prims.empty()) return false;
  int lights[64];
  int nl = 0;
  for (size_t i = 0; i < s.prims.size() && nl < 64; ++i)
    if (s.prims[i].has_em) lights[nl++] = (int)i;
  if (nl == 0) return false;
  int idx = (int)(rng.nextI() % (uint32_t)nl);
  if (chosen) *chosen = lights[idx];
  const Prim& pr = s.prims[lights[idx]];
  if (pr.type == MUPSI_PRIM_SPHERE) return sphere_sample_direct(pr, p, rng, ls);
  return mesh_sample_direct(pr, s.meshes[pr.mesh], p, rng, ls);
}

static const float kDefaultAlbedo = 0.8f;  // bsdf.cpp:11

static V3 albedo_of(const Prim* pr, float u, float v) {
  if (pr && pr->albedo.present()) return pr->albedo.at(u, v);
  return mk(kDefaultAlbedo, kDefaultAlbedo, kDefaultAlbedo);
}

// Bsdf::weight = eval / pdf for the light direction (bsdf.h:51-55, bsdf.cpp:14-21,39-50,57-74)
static V3 bsdf_weight(int bsdf, V3 alb, V3 wo, V3 n, V3 wi) {
  if (bsdf == MUPSI_BSDF_LAMBERTIAN) {
    float cosv = dot(wi, n);
    float pdf = fdiv_pi(std::max(cosv, 0.0f));
    if (pdf == 0.0f) return mk(0, 0, 0);
    const float pif = (float)M_PI;
    V3 ev = mk(alb.x * cosv / pif, alb.y * cosv / pif, alb.z * cosv / pif);
    return divs(ev, pdf);
  }
  if (bsdf == MUPSI_BSDF_NULL) {
    return dot(wi, neg(wo)) > 0.999f ? mk(1, 1, 1) : mk(0, 0, 0);
  }
  float k = 2.0f * dot(wi, n);
  V3 wr = mk(fmaf(k, n.x, -wi.x), fmaf(k, n.y, -wi.y), fmaf(k, n.z, -wi.z));
  return dot(wr, wo) > 0.999f ? mk(1, 1, 1) : mk(0, 0, 0);
}

// Bsdf::sample: fills wi and weight (bsdf.cpp:23-37,51-55,75-79)
static void bsdf_sample(int bsdf, V3 alb, V3 wo, V3 n, Pcg& rng, V3& wi, V3& weight) {
  if (bsdf == MUPSI_BSDF_LAMBERTIAN) {
    float r1 = rng.next1D();
    float r2 = rng.next1D();
    float phi = (float)(2 * M_PI * (double)r1);
    float cosT = (float)std::sqrt((double)(1 - r2));
    float sinT = (float)std::sqrt((double)r2);
    V3 lw = mk((float)((double)sinT * std::cos((double)phi)),
               (float)((double)sinT * std::sin((double)phi)), cosT);
    wi = to_global(frame(n), lw);
    weight = alb;
  } else if (bsdf == MUPSI_BSDF_NULL) {
    wi = neg(wo);
    weight = mk(1, 1, 1);
  } else {
    float k = 2.0f * dot(wo, n);
    wi = mk(fmaf(k, n.x, -wo.x), fmaf(k, n.y, -wo.y), fmaf(k, n.z, -wo.z));
    weight = mk(1, 1, 1);
  }
}

// Skydrome::intersect + BitmapTexture lookup (skydrome.cpp:9-25, texture.cpp:52-69)
static V3 sky_lookup(const Scene& s, V3 d, int* texel) {
  float u = (float)(((double)atan2f(d.z, d.x) + M_PI) / (2.0f * M_PI));
  float v = (float)((double)acosf(std::clamp(-d.y, -1.0f, 1.0f)) / M_PI);
  return s.sky.at(u, v, texel);
}

// The decision margin of one march of a traced path, for *amb of trace(): the
// same march in FP64 at the same FP32 t-grid points (march(..., fp64, track)),
// whose amb is the smallest |f| / (1 + |mu| + l1) at a visited grid point; 0 when
// the FP64 march decides otherwise than the FP32 one (the FP32 decision then
// rests on rounding).  GP media only (the SDF and grid marches have no margin).
static void track_margin(const Scene& s, int mode, const Ray& ray, uint32_t seed, const Cond& c,
                         const MarchOut& mo, double* amb) {
  if (!amb || s.medium_kind != MUPSI_MEDIUM_GP) return;
  const MarchOut m64 = march(s, mode, ray, seed, c, true, true);
  double a = m64.amb;
  if (m64.hit != mo.hit || (mo.hit && m64.t != mo.t)) a = 0.0;
  *amb = std::min(*amb, a);
}

// PathTracer::trace (trace.cpp:118-213), handleSurface (:32-75), handleVolume (:77-116).
// amb (may be null): receives the smallest decision margin of the path's
// marches (track_margin), +inf for a path with no GP-medium march.
static V3 trace(const Scene& s, int x, int y, int k, uint32_t seed, double* amb = nullptr) {
  if (amb) *amb = INFINITY;
  uint32_t pix_seed = make_seed4((uint32_t)x, (uint32_t)y, (uint32_t)k, seed);
  Pcg rng = pcg_ctor(pix_seed);
  Ray ray = make_ray(s.cam.pos, cam_dir(s.cam, x, y));
  V3 L = mk(0, 0, 0), T = mk(1, 1, 1);
  int bounce = 0;
  bool medium = false;
  Cond cond;
  bool hasHit;
  const bool single = s.gp_mode != MUPSI_GP_RENEWAL_PLUS;
  do {
    hasHit = false;
    bool exited = true;
    if (medium && !s.skip_medium) {
      uint32_t mseed = single ? seed : make_seed5(x, y, k, seed, (uint32_t)bounce);
      uint32_t iseed = constant_sampler_i(mseed);
      MarchOut mo = march(s, 0, ray, iseed, cond, false, false);
      track_margin(s, 0, ray, iseed, cond, mo, amb);
      if (mo.hit) {
        V3 P = mo.p, n = mo.n, wo = neg(ray.d);
        V3 alb = mk(kDefaultAlbedo, kDefaultAlbedo, kDefaultAlbedo);
        LightSample ls;
        if (choose_light(s, P, rng, ls)) {
          Ray sr = make_ray(madd(n, s.eps, P), ls.d);
          float bias = ls.dist - s.eps * std::fabs(dot(ls.d, n));
          sr.farT = bias * 0.99f;
          MarchOut tr = march(s, 1, sr, iseed, cond, false, false);
          track_margin(s, 1, sr, iseed, cond, tr, amb);
          float trans = tr.hit ? 0.0f : 1.0f;
          V3 f = bsdf_weight(MUPSI_BSDF_LAMBERTIAN, alb, wo, n, ls.d);
          float ac = std::fabs(dot(ls.d, n));
          L = mk(fmaf(trans * (T.x * (f.x * ls.w.x)), ac, L.x),
                 fmaf(trans * (T.y * (f.y * ls.w.y)), ac, L.y),
                 fmaf(trans * (T.z * (f.z * ls.w.z)), ac, L.z));
        }
        V3 wi, wt;
        bsdf_sample(MUPSI_BSDF_LAMBERTIAN, alb, wo, n, rng, wi, wt);
        T = cmul(T, wt);
        float surv = max3(T);
        if (rng.next1D() > surv) break;
        T = divs(T, surv);
        uint32_t nseed = single ? seed : make_seed5(x, y, k, seed, (uint32_t)(bounce + 1));
        if (s.medium_kind == MUPSI_MEDIUM_GP)  // only GPMedium re-conditions (trace.cpp:157-160)
          cond = condition_f32(s, P, n, constant_sampler_i(nseed));
        ray = make_ray(madd(n, s.eps, P), wi);
        ray.nearT = s.near_eps;
        ++bounce;
        exited = false;
        hasHit = true;
      }
    }
    if (exited) {
      V3 P, Ng;
      float uu = 0, vv = 0;
      int hp = scene_intersect(s, ray, P, Ng, uu, vv);
      if (hp >= 0) {
        const Prim& pr = s.prims[hp];
        V3 wo = neg(ray.d), n = Ng;
        V3 alb = pr.bsdf == MUPSI_BSDF_LAMBERTIAN ? albedo_of(&pr, uu, vv) : mk(1, 1, 1);
        LightSample ls;
        if (choose_light(s, P, rng, ls)) {
          Ray sr = make_ray(madd(n, s.eps, P), ls.d);
          float bias = ls.dist - s.eps * std::fabs(dot(ls.d, n));
          sr.farT = bias * 0.99f;
          if (!scene_occluded(s, sr)) {
            V3 a = bsdf_weight(pr.bsdf, alb, wo, n, ls.d);
            float ac = std::fabs(dot(ls.d, n));
            L = mk(fmaf(T.x * (a.x * ls.w.x), ac, L.x), fmaf(T.y * (a.y * ls.w.y), ac, L.y),
                   fmaf(T.z * (a.z * ls.w.z), ac, L.z));
          }
        }
        V3 wi, wt;
        bsdf_sample(pr.bsdf, alb, wo, n, rng, wi, wt);
        if (pr.has_em) {
          float m = std::max(0.0f, dot(n, wi));
          L = mk(fmaf(T.x, pr.em.x * m, L.x), fmaf(T.y, pr.em.y * m, L.y),
                 fmaf(T.z, pr.em.z * m, L.z));
        }
        T = cmul(T, wt);
        float surv = max3(T);
        if (rng.next1D() > surv) break;
        T = divs(T, surv);
        bool back = dot(wi, n) < 0;
        float bs = back ? -1.0f : 1.0f;
        medium = back ? (pr.int_medium && s.has_gp) : false;
        V3 bn = mul(Ng, bs);
        ray = make_ray(madd(bn, s.eps, P), wi);
        ray.nearT = s.near_eps;
        ++bounce;
        hasHit = true;
      } else if (s.sky.present()) {
        V3 e = sky_lookup(s, ray.d, nullptr);
        L = mk(fmaf(T.x, e.x, L.x), fmaf(T.y, e.y, L.y), fmaf(T.z, e.z, L.z));
      }
    }
  } while (hasHit && bounce < s.max_bounce);
  return L;
}

static Scene* build_scene(const mupsi_scene_desc& d) {
  Scene* s = new Scene();
  s->cam = make_cam(d);
  s->max_bounce = d.max_bounce;
  s->gp_mode = d.gp_mode;
  s->skip_medium = d.skip_medium;
  s->has_gp = d.has_gp_medium != 0;
  s->medium_kind = s->has_gp ? d.medium_kind : MUPSI_MEDIUM_GP;
  s->mean_type = d.mean_type;
  s->mean_c = mk(d.mean_center[0], d.mean_center[1], d.mean_center[2]);
  s->mean_r = d.mean_radius;
  if (s->has_gp && s->medium_kind != MUPSI_MEDIUM_GRID && d.mean_type == MUPSI_MEAN_MESH)
    s->mean_mesh.init(d.mean_mesh);
  if (s->medium_kind == MUPSI_MEDIUM_GP) s->gp = make_gp(d.kernel);
  if (s->has_gp && s->medium_kind == MUPSI_MEDIUM_GRID && d.grid.mean && d.grid.resolution >= 2) {
    // GridMeanFunction ctor (meanfunction.cpp:41-54): inv cell = res / (max - min)
    Grid& g = s->grid;
    g.res = d.grid.resolution;
    g.bmin = mk(d.grid.bounds_min[0], d.grid.bounds_min[1], d.grid.bounds_min[2]);
    g.inv = mk((float)g.res / (d.grid.bounds_max[0] - d.grid.bounds_min[0]),
               (float)g.res / (d.grid.bounds_max[1] - d.grid.bounds_min[1]),
               (float)g.res / (d.grid.bounds_max[2] - d.grid.bounds_min[2]));
    g.noise_scale = d.grid.noise_scale;
    const size_t n = (size_t)g.res * g.res * g.res;
    g.mean.assign(d.grid.mean, d.grid.mean + n);
    if (d.grid.variance) g.var.assign(d.grid.variance, d.grid.variance + n);
  }
  s->sky = make_tex(d.sky);
  s->meshes.reserve(d.n_primitives);
  for (int i = 0; i < d.n_primitives; ++i) {
    const mupsi_primitive_desc& pd = d.primitives[i];
    Prim pr;
    pr.type = pd.type;
    pr.bsdf = pd.bsdf;
    // emission: the reference parses it for spheres only (controller.cpp:225-228);
    // a mesh given one behaves as Mesh::getEmission / sampleDirect would
    pr.has_em = pd.has_emission != 0;
    pr.em = mk(pd.emission[0], pd.emission[1], pd.emission[2]);
    pr.int_medium = pd.type == MUPSI_PRIM_SPHERE && pd.int_medium;
    if (pd.type == MUPSI_PRIM_SPHERE) {
      pr.xf = make_xf(mk(pd.center[0], pd.center[1], pd.center[2]), pd.radius);
    } else {
      s->meshes.emplace_back();
      s->meshes.back().init(pd.mesh);
      pr.mesh = (int)s->meshes.size() - 1;
      if (pd.bsdf == MUPSI_BSDF_LAMBERTIAN) pr.albedo = make_tex(pd.albedo);
    }
    s->prims.push_back(std::move(pr));
  }
  return s;
}

static Cond cond_from(const float* c, int64_t i) {
  Cond r;
  if (!c) return r;
  const float* q = c + 8 * i;
  r.active = q[0] != 0.0f;
  r.C = mk(q[1], q[2], q[3]);
  r.u = q[4];
  r.g = mk(q[5], q[6], q[7]);
  return r;
}
static const Mesh* mesh_of(const Scene* s, int prim) {
  if (prim < 0) return &s->mean_mesh;
  if (prim >= (int)s->prims.size() || s->prims[prim].type != MUPSI_PRIM_MESH) return nullptr;
  return &s->meshes[s->prims[prim].mesh];
}
static Ray ray_from(const float* r, int64_t i) {
  const float* q = r + 8 * i;
  Ray ray = make_ray(mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]));
  ray.nearT = q[6];
  ray.farT = q[7];
  return ray;
}

}  // namespace orc

using namespace orc;

extern "C" {

uint32_t orc_make_seed(const uint32_t* vals, int32_t n) {
  uint32_t h = 0;
  for (int i = 0; i < n; ++i) h = hash_combine(h, vals[i]);
  return h;
}
void orc_pcg_draws(uint64_t seed, int32_t ctor, int32_t n, uint32_t* out) {
  Pcg r = ctor ? pcg_ctor(seed) : Pcg{seed};
  for (int i = 0; i < n; ++i) out[i] = r.nextI();
}
float orc_normalized_uint(uint32_t i) { return Pcg::norm_uint(i); }
uint32_t orc_constant_sampler_i(uint32_t seed) { return constant_sampler_i(seed); }
float orc_psi_var(const mupsi_gp_params* gp) { return make_gp(*gp).var; }

void orc_cell_impulses(const mupsi_gp_params* gp, const int32_t* cell3, const uint32_t* seed,
                       int64_t n, uint32_t* cell_seed, float* pos, int8_t* w) {
  GP g = make_gp(*gp);
  Impulse imp[64];
  for (int64_t i = 0; i < n; ++i) {
    cell_impulses(g, cell3[3 * i], cell3[3 * i + 1], cell3[3 * i + 2], seed[i], &cell_seed[i], imp);
    for (int j = 0; j < g.ppc; ++j) {
      pos[(i * g.ppc + j) * 3 + 0] = imp[j].s.x;
      pos[(i * g.ppc + j) * 3 + 1] = imp[j].s.y;
      pos[(i * g.ppc + j) * 3 + 2] = imp[j].s.z;
      w[i * g.ppc + j] = (int8_t)imp[j].w;
    }
  }
}

void orc_psi_eval(const mupsi_gp_params* gp, const float* p3, const uint32_t* seed, int64_t n,
                  int32_t fp64, double* out4, double* l1) {
  GP g = make_gp(*gp);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    V3 p = mk(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
    if (fp64) {
      double m;
      F4d f = raw_noise_f64(g, p, seed[i], &m);
      out4[4 * i] = f.v;
      out4[4 * i + 1] = f.gx;
      out4[4 * i + 2] = f.gy;
      out4[4 * i + 3] = f.gz;
      if (l1) l1[i] = m;
    } else {
      float m;
      F4 f = raw_noise_f32(g, p, seed[i], &m);
      out4[4 * i] = f.v;
      out4[4 * i + 1] = f.gx;
      out4[4 * i + 2] = f.gy;
      out4[4 * i + 3] = f.gz;
      if (l1) l1[i] = m;
    }
  }
}

void* orc_scene_create(const mupsi_scene_desc* desc) { return build_scene(*desc); }
void orc_scene_free(void* s) { delete (Scene*)s; }

void orc_hash11(const float* p3, const uint32_t* seed, int64_t n, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = hash11(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2], seed[i]);
}

void orc_field_eval(void* sc, const float* p3, const uint32_t* seed, const float* cond,
                    int64_t n, int32_t fp64, double* out4) {
  const Scene& s = *(Scene*)sc;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    V3 p = mk(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
    Cond c = cond_from(cond, i);
    if (s.medium_kind != MUPSI_MEDIUM_GP) {  // SDF: the mean; grid medium: grids + noise
      F4 f = s.medium_kind == MUPSI_MEDIUM_SDF ? F4{mean_eval(s, p), 0, 0, 0}
                                               : grid_medium(s, p, seed[i]);
      if (s.medium_kind == MUPSI_MEDIUM_SDF) {
        V3 g = mean_grad(s, p);
        f.gx = g.x;
        f.gy = g.y;
        f.gz = g.z;
      }
      out4[4 * i] = f.v;
      out4[4 * i + 1] = f.gx;
      out4[4 * i + 2] = f.gy;
      out4[4 * i + 3] = f.gz;
    } else if (fp64) {
      F4d f = field_f64(s, p, seed[i], c);
      out4[4 * i] = f.v;
      out4[4 * i + 1] = f.gx;
      out4[4 * i + 2] = f.gy;
      out4[4 * i + 3] = f.gz;
    } else {
      F4 f = field_f32(s, p, seed[i], c);
      out4[4 * i] = f.v;
      out4[4 * i + 1] = f.gx;
      out4[4 * i + 2] = f.gy;
      out4[4 * i + 3] = f.gz;
    }
  }
}

void orc_march(void* sc, int32_t mode, const float* rays, const uint32_t* seed, const float* cond,
               int64_t n, int32_t fp64, uint8_t* hit, float* t, float* p3, float* n3,
               int32_t* steps, double* amb, double* cell) {
  const Scene& s = *(Scene*)sc;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n; ++i) {
    Ray r = ray_from(rays, i);
    MarchOut m = march(s, mode, r, seed[i], cond_from(cond, i), fp64 != 0, amb || cell);
    hit[i] = m.hit;
    t[i] = m.t;
    p3[3 * i] = m.p.x;
    p3[3 * i + 1] = m.p.y;
    p3[3 * i + 2] = m.p.z;
    n3[3 * i] = m.n.x;
    n3[3 * i + 1] = m.n.y;
    n3[3 * i + 2] = m.n.z;
    steps[i] = m.steps;
    if (amb) amb[i] = m.amb;
    if (cell) cell[i] = m.cell;
  }
}

void orc_condition(void* sc, const float* c3, const float* n3, const uint32_t* seed_next,
                   int64_t n, int32_t fp64, double* u, double* g3) {
  const Scene& s = *(Scene*)sc;
  for (int64_t i = 0; i < n; ++i) {
    V3 C = mk(c3[3 * i], c3[3 * i + 1], c3[3 * i + 2]);
    V3 nn = mk(n3[3 * i], n3[3 * i + 1], n3[3 * i + 2]);
    if (fp64) {
      condition_f64(s, C, nn, seed_next[i], &u[i], &g3[3 * i]);
    } else {
      Cond c = condition_f32(s, C, nn, seed_next[i]);
      u[i] = c.u;
      g3[3 * i] = c.g.x;
      g3[3 * i + 1] = c.g.y;
      g3[3 * i + 2] = c.g.z;
    }
  }
}

void orc_mean_eval(void* sc, const float* p3, int64_t n, double* out4) {
  const Scene& s = *(Scene*)sc;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    V3 p = mk(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
    float mu = mean_eval(s, p);
    V3 g = mean_grad(s, p);
    out4[4 * i] = mu;
    out4[4 * i + 1] = g.x;
    out4[4 * i + 2] = g.y;
    out4[4 * i + 3] = g.z;
  }
}

int32_t orc_bvh_build(const float* vertices, int32_t n_vertices, const int32_t* faces,
                      int32_t n_faces, float* nodes_out, uint32_t* tri_index_out,
                      uint32_t* root_out) {
  std::vector<V3> v(n_vertices);
  for (int i = 0; i < n_vertices; ++i)
    v[i] = mk(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]);
  std::vector<Tri> f(n_faces);
  for (int i = 0; i < n_faces; ++i) f[i] = Tri{faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]};
  BVH b;
  b.build(v.data(), f.data(), (uint32_t)n_faces);
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    const Node& nd = b.nodes[i];
    float* o = nodes_out + 8 * i;
    for (int a = 0; a < 3; ++a) {
      o[a] = nd.bmin[a];
      o[3 + a] = nd.bmax[a];
    }
    std::memcpy(o + 6, &nd.left, 4);
    std::memcpy(o + 7, &nd.right, 4);
  }
  for (size_t i = 0; i < b.tri.size(); ++i) tri_index_out[i] = b.tri[i];
  *root_out = b.root;
  return (int32_t)b.nodes.size();
}

void orc_mesh_intersect(void* sc, int32_t prim, const float* rays, int64_t n, int32_t* tri,
                        float* t, float* u, float* v) {
  const Mesh* m = mesh_of((Scene*)sc, prim);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    Ray r = ray_from(rays, i);
    uint32_t ti = 0;
    float uu = 0, vv = 0;
    bool h = m && m->intersect(r, ti, uu, vv);
    tri[i] = h ? (int32_t)ti : -1;
    t[i] = h ? r.farT : 0.0f;
    u[i] = uu;
    v[i] = vv;
  }
}

// Brute force over every triangle in world space (BVH::intersectBruteForce, bvh.cpp:357-374,
// applied through Mesh::intersect's local-space transform).
void orc_mesh_intersect_brute(void* sc, int32_t prim, const float* rays, int64_t n,
                              int32_t* tri, float* t) {
  const Mesh* m = mesh_of((Scene*)sc, prim);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    Ray ray = ray_from(rays, i);
    Ray lr = ray;
    float tcorr = 1.0f;
    if (m && !m->xf.identity) {
      V3 lo = xf_to_local(m->xf, ray.o);
      V3 ld = xf_dir_to_local(m->xf, ray.d);
      float len = norm(ld);
      tcorr = norm(xf_dir_to_world(m->xf, divs(ld, len)));
      lr = make_ray(lo, ld);
      lr.nearT = ray.nearT * len;
      lr.farT = ray.farT * len;
    }
    int32_t best = -1;
    if (m) {
      for (size_t k = 0; k < m->faces.size(); ++k) {
        float tt, uu, vv;
        const Tri& f = m->faces[k];
        if (ray_tri(lr, m->verts[f.a], m->verts[f.b], m->verts[f.c], tt, uu, vv)) {
          lr.farT = tt;
          best = (int32_t)k;
        }
      }
    }
    tri[i] = best;
    t[i] = best >= 0 ? lr.farT * tcorr : 0.0f;
  }
}

void orc_mesh_occluded(void* sc, int32_t prim, const float* rays, int64_t n, uint8_t* occ) {
  const Mesh* m = mesh_of((Scene*)sc, prim);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) occ[i] = m && m->occluded(ray_from(rays, i));
}

void orc_mesh_closest(void* sc, int32_t prim, const float* p3, int64_t n, float* dist,
                      int32_t* tri, float* closest3) {
  const Mesh* m = mesh_of((Scene*)sc, prim);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    uint32_t ti = 0;
    V3 o = mk(0, 0, 0);
    float d = m ? m->dist_to_point(mk(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]), ti, o) : 0.0f;
    dist[i] = d;
    tri[i] = (int32_t)ti;
    closest3[3 * i] = o.x;
    closest3[3 * i + 1] = o.y;
    closest3[3 * i + 2] = o.z;
  }
}

void orc_mesh_inside(void* sc, int32_t prim, const float* p3, int64_t n, uint8_t* inside) {
  const Mesh* m = mesh_of((Scene*)sc, prim);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i)
    inside[i] = m && m->inside(mk(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]));
}

// Traversal cost per point of distToPoint (kind 0) or inside (kind 1).
void orc_mesh_query_cost(void* sc, int32_t prim, int32_t kind, const float* p3, int64_t n,
                         uint32_t* nodes, uint32_t* tris) {
  const Mesh* m = mesh_of((Scene*)sc, prim);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    TravCost c;
    const V3 p = mk(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
    if (m) {
      if (kind == 0) {
        uint32_t ti = 0;
        V3 o = mk(0, 0, 0);
        (void)m->dist_to_point(p, ti, o, &c);
      } else {
        (void)m->inside(p, &c);
      }
    }
    nodes[i] = (uint32_t)c.nodes;
    tris[i] = (uint32_t)c.tris;
  }
}

void orc_camera_rays(void* sc, const int32_t* x, const int32_t* y, int64_t n, float* dir3) {
  const Scene& s = *(Scene*)sc;
  for (int64_t i = 0; i < n; ++i) {
    V3 d = make_ray(s.cam.pos, cam_dir(s.cam, x[i], y[i])).d;
    dir3[3 * i] = d.x;
    dir3[3 * i + 1] = d.y;
    dir3[3 * i + 2] = d.z;
  }
}

void orc_sky_lookup(void* sc, const float* dir3, int64_t n, float* rgb, int32_t* texel) {
  const Scene& s = *(Scene*)sc;
  for (int64_t i = 0; i < n; ++i) {
    int tx = -1;
    V3 c = s.sky.present() ? sky_lookup(s, mk(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]), &tx)
                           : mk(0, 0, 0);
    rgb[3 * i] = c.x;
    rgb[3 * i + 1] = c.y;
    rgb[3 * i + 2] = c.z;
    texel[i] = tx;
  }
}

void orc_trace_paths(void* sc, const int32_t* x, const int32_t* y, const int32_t* k, int64_t n,
                     float* rgb, int32_t nthreads) {
  const Scene& s = *(Scene*)sc;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n; ++i) {
    V3 c = trace(s, x[i], y[i], k[i], 42u);
    rgb[3 * i] = c.x;
    rgb[3 * i + 1] = c.y;
    rgb[3 * i + 2] = c.z;
  }
}

void orc_trace_paths_amb(void* sc, const int32_t* x, const int32_t* y, const int32_t* k,
                         int64_t n, float* rgb, double* amb, int32_t nthreads) {
  const Scene& s = *(Scene*)sc;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n; ++i) {
    V3 c = trace(s, x[i], y[i], k[i], 42u, &amb[i]);
    rgb[3 * i] = c.x;
    rgb[3 * i + 1] = c.y;
    rgb[3 * i + 2] = c.z;
  }
}

void orc_choose_light(void* sc, const float* p3, uint64_t* state, int64_t n, int32_t* light,
                      float* w3, float* d3, float* dist, float* pdf) {
  const Scene& s = *(Scene*)sc;
  for (int64_t i = 0; i < n; ++i) {
    Pcg rng{state[i]};
    LightSample ls{};
    int chosen = -1;
    const bool ok = choose_light(s, mk(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]), rng, ls, &chosen);
    state[i] = rng.s;
    light[i] = ok ? chosen : -1 - (chosen >= 0 ? 1 : 0);
    w3[3 * i] = ls.w.x;
    w3[3 * i + 1] = ls.w.y;
    w3[3 * i + 2] = ls.w.z;
    d3[3 * i] = ls.d.x;
    d3[3 * i + 1] = ls.d.y;
    d3[3 * i + 2] = ls.d.z;
    dist[i] = ls.dist;
    pdf[i] = ls.pdf;
  }
}

double orc_render(void* sc, int32_t k_begin, int32_t k_count, int32_t pixel_stride,
                  int32_t pixel_offset, int32_t nthreads, double* sums, int64_t* paths_out) {
  const Scene& s = *(Scene*)sc;
  const int w = s.cam.w, h = s.cam.h;
  if (pixel_stride < 1) pixel_stride = 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int64_t paths = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (int k = k_begin; k < k_begin + k_count; ++k) {
    // Renderer::startRenderProgressive: OpenMP collapse(2), static schedule (renderer.cpp:28-46)
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : paths)
    for (int j = 0; j < h; ++j)
      for (int i = 0; i < w; ++i) {
        int64_t lin = (int64_t)j * w + i;
        if ((lin % pixel_stride) != pixel_offset) continue;
        V3 e = trace(s, i, j, k, 42u);
        sums[3 * lin] += e.x;
        sums[3 * lin + 1] += e.y;
        sums[3 * lin + 2] += e.z;
        ++paths;
      }
  }
  auto t1 = std::chrono::steady_clock::now();
  if (paths_out) *paths_out = paths;
  return std::chrono::duration<double>(t1 - t0).count();
}

int32_t orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
