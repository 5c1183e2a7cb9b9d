// kernels.hip — wavefront GPIS path tracer for gfx950 and its C ABI.
//
// Data flow of one mupsi_render_passes launch (paths = pixels x passes):
//   k_init   : camera ray + path RNG for every path; all paths enter queue A.
//   per bounce iteration (<= max_bounce):
//     k_march  : persistent, lane-per-ray GP-medium marcher over queue M
//                (paths whose ray is inside the GP medium).  Lanes refill from
//                the queue as their rays finish, so uneven march lengths do
//                not idle the wave.
//     k_shade  : medium hits (normal, NEE, Lambertian sample, RR, Renewal+
//                FP64 solve) and surface/sky events for queue A; survivors are
//                compacted into the next A and M queues with wave ballots.
//   k_accum  : per-pixel FP64 sum of the path radiances (fixed pass order).
// Path state is SoA in HBM, one slot per path of the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mupsi.h"
#include "device.h"
#include "guide_host.h"
#include "host_bvh.h"
#include "march_cached.h"
#include "runtime.h"
#include "scene_host.h"

using namespace mupsi;
using namespace mupsi::dev;

// ====================================================================== device: path state

constexpr uint32_t kBounceMask = 0xffffu;
constexpr uint32_t kFlagMedium = 1u << 16;
constexpr uint32_t kFlagCond = 1u << 17;
constexpr uint32_t kRecShift = 18;  // guided vertices recorded for training (3 bits)
constexpr uint32_t kRecMask = 7u << kRecShift;
constexpr uint32_t kFlagOpen = 1u << 21;  // the last record still collects its direct radiance
constexpr uint32_t kFlagNeeRec = 1u << 22;  // the record of this event waits for its light sample

struct PathSoA {
  uint64_t* rng;
  float *ox, *oy, *oz, *dx, *dy, *dz, *nt;
  float *Tx, *Ty, *Tz, *Lx, *Ly, *Lz;
  float* mis;  // guided tracer: product of the MIS factors (cos/pi)/q of the path
  uint32_t* flags;
  float *cx, *cy, *cz, *cu, *cgx, *cgy, *cgz;
  uint32_t* mseed;
  uint8_t* mhit;
  float *mt, *mpx, *mpy, *mpz;
  uint16_t* sid;  // batched launches: the path's scene
};

// Deferred light samples of GP-medium hits (handleVolume's NEE, trace.cpp:84-98):
// the shadow ray, the realization and conditioning it is marched in, its march
// result, and the contribution waiting for it: L = fmaf(pend, ac, L) if the
// march finds no sign change (transmittance 1), as the reference adds it.
struct ShadowSoA {
  float *ox, *oy, *oz, *dx, *dy, *dz, *nt, *ft;
  uint32_t* seed;
  uint8_t* active;
  float *cx, *cy, *cz, *cu, *cgx, *cgy, *cgz;
  uint8_t* hit;
  float *t, *px, *py, *pz;
  float *pend_x, *pend_y, *pend_z, *ac;
};

// A batched launch (several scenes in one wavefront): scene s owns the path ids
// [off, off + kc * npix), passes k_begin .. k_begin + kc - 1 of its W-wide image,
// pass-major.  The plan lives in device memory so that one captured graph can
// be replayed with the next launch's plan.
struct LaunchPlan {
  int64_t off;
  int npix, W, k_begin, kc;
};

// Path id -> (x, y, k): explicit lists (replay), the pixel grid of a launch, or a
// batched launch's plan (with the path's scene id).
struct PathSrc {
  int W, npix, k_begin;
  const int* k_begin_dev;   // graph replays: the first pass, read from device memory
  const int32_t *xs, *ys, *ks;
  const LaunchPlan* plan;   // batched launch, or null
  int nplan;
  const int64_t* total;     // batched launch: paths in this launch (device)
};
__device__ __forceinline__ void plan_xyk(const LaunchPlan& L, int64_t pid, int& x, int& y, int& k) {
  const int64_t local = pid - L.off;
  const int pix = (int)(local % L.npix);
  x = pix % L.W;
  y = pix / L.W;
  k = L.k_begin + (int)(local / L.npix);
}
__device__ __forceinline__ void path_xyk(const PathSrc& s, int pid, int& x, int& y, int& k,
                                         const uint16_t* sid = nullptr) {
  if (s.plan) {
    plan_xyk(s.plan[sid[pid]], pid, x, y, k);
  } else if (s.xs) {
    x = s.xs[pid];
    y = s.ys[pid];
    k = s.ks[pid];
  } else {
    const int pix = pid % s.npix;
    x = pix % s.W;
    y = pix / s.W;
    k = (s.k_begin_dev ? *s.k_begin_dev : s.k_begin) + pid / s.npix;
  }
}

// The scene of a path: the kernel argument in a single-scene launch, the
// per-scene table in a batched one.
template <bool kMulti>
__device__ __forceinline__ const DevScene& pick_scene(const DevScene& S0, const DevScene* scenes,
                                                      const uint16_t* sid, int pid) {
  if constexpr (kMulti) return scenes[sid[pid]];
  else return S0;
}

// Slot of this lane among the wave's lanes with pred, after one atomic per wave
// on count (every lane of the wave must call it); -1 when pred is false.
__device__ __forceinline__ int wave_slot(bool pred, int* count) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return -1;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader);
  return pred ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// Append value to queue if pred; one atomic per wave.
__device__ __forceinline__ void wave_push(bool pred, int value, int* queue, int* count) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader);
  if (pred) queue[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  return v;
}

// ====================================================================== device: camera, sky, BSDF

// Camera::generateRay (camera.cpp:5-19); tan(fov/2) is precomputed on the host
// with the reference's degrees-as-radians argument.
__device__ __forceinline__ F3 camera_dir(const DevScene& S, int x, int y) {
  const float cxf = (x + 0.5f) / S.w;
  const float cyf = (y + 0.5f) / S.h;
  const float nx = fmaf(2.0f, cxf, -1.0f) * S.aspect * S.tan_half;
  const float ny = fmaf(2.0f, cyf, -1.0f) * S.tan_half;
  const F3 rc = normalized(mk(nx, ny, 1.0f));
  const F3 w = mk(fmaf(S.cam_dir[0], rc.z, fmaf(S.cam_tup[0], rc.y, S.cam_right[0] * rc.x)),
                  fmaf(S.cam_dir[1], rc.z, fmaf(S.cam_tup[1], rc.y, S.cam_right[1] * rc.x)),
                  fmaf(S.cam_dir[2], rc.z, fmaf(S.cam_tup[2], rc.y, S.cam_right[2] * rc.x)));
  return normalized(w);
}

// BitmapTexture::operator[] (texture.cpp:52-69)
__device__ __forceinline__ F3 tex_at(const DevTex& t, float u, float v, int* texel) {
  u = u - floorf(u);
  v = v - floorf(v);
  v = 1.0f - v;
  int x = (int)(u * t.w) % t.w;
  int y = (int)(v * t.h) % t.h;
  if (x < 0) x += t.w;
  if (y < 0) y += t.h;
  const int i = y * t.w + x;
  if (texel) *texel = i;
  return mk(t.rgb[3 * i], t.rgb[3 * i + 1], t.rgb[3 * i + 2]);
}
// Skydrome::intersect (skydrome.cpp:9-18): the float library calls are taken in
// double and rounded once, which reproduces the correctly rounded results the
// host math library returns.
__device__ __forceinline__ F3 sky_lookup(const DevScene& S, F3 d, int* texel) {
  const float at = (float)atan2((double)d.z, (double)d.x);
  const float u = (float)(((double)at + M_PI) / (2.0f * M_PI));
  const float cl = fminf(fmaxf(-d.y, -1.0f), 1.0f);
  const float ac = (float)acos((double)cl);
  const float v = (float)((double)ac / M_PI);
  return tex_at(S.sky, u, v, texel);
}

struct Frame {
  F3 n, t, b;
};
// TangentFrame(n), Duff et al. 2017 (sample.h:25-35)
__device__ __forceinline__ Frame make_frame(F3 nin) {
  Frame f;
  f.n = normalized(nin);
  const float sign = copysignf(1.0f, f.n.z);
  const float a = -1.0f / (sign + f.n.z);
  const float b = f.n.x * f.n.y * a;
  f.t = normalized(mk(fmaf(sign * f.n.x * f.n.x, a, 1.0f), sign * b, -sign * f.n.x));
  f.b = normalized(mk(b, fmaf(f.n.y * f.n.y, a, sign), -f.n.y));
  return f;
}
__device__ __forceinline__ F3 to_global(const Frame& f, F3 p) {
  return mk(fmaf(f.n.x, p.z, fmaf(f.b.x, p.y, f.t.x * p.x)),
            fmaf(f.n.y, p.z, fmaf(f.b.y, p.y, f.t.y * p.x)),
            fmaf(f.n.z, p.z, fmaf(f.b.z, p.y, f.t.z * p.x)));
}

// Bsdf::weight for a light direction (bsdf.h:51-55, bsdf.cpp:14-21,39-50,57-74)
__device__ F3 bsdf_weight(int bsdf, F3 alb, F3 wo, F3 n, F3 wi) {
  if (bsdf == MUPSI_BSDF_LAMBERTIAN) {
    const float cosv = dot(wi, n);
    const float pdf = (float)((double)fmaxf(cosv, 0.0f) / M_PI);
    if (pdf == 0.0f) return mk(0, 0, 0);
    const float pif = (float)M_PI;
    return divs(mk(alb.x * cosv / pif, alb.y * cosv / pif, alb.z * cosv / pif), pdf);
  }
  if (bsdf == MUPSI_BSDF_NULL) return dot(wi, neg(wo)) > 0.999f ? mk(1, 1, 1) : mk(0, 0, 0);
  const float k = 2.0f * dot(wi, n);
  const F3 wr = mk(fmaf(k, n.x, -wi.x), fmaf(k, n.y, -wi.y), fmaf(k, n.z, -wi.z));
  return dot(wr, wo) > 0.999f ? mk(1, 1, 1) : mk(0, 0, 0);
}
// Bsdf::sample (bsdf.cpp:23-37,51-55,75-79).  The Lambertian sample takes
// phi, sqrt, cos and sin in double as the reference's unqualified calls do.
__device__ __forceinline__ F3 lambert_dir(F3 n, float r1, float r2) {
  const double phi = (double)(float)(2 * M_PI * (double)r1);
  const float cosT = (float)sqrt((double)(1 - r2));
  const float sinT = (float)sqrt((double)r2);
  double sp, cp;
  sincos(phi, &sp, &cp);
  const F3 lw = mk((float)((double)sinT * cp), (float)((double)sinT * sp), cosT);
  return to_global(make_frame(n), lw);
}
__device__ void bsdf_sample(int bsdf, F3 alb, F3 wo, F3 n, Pcg& rng, F3& wi, F3& weight) {
  if (bsdf == MUPSI_BSDF_LAMBERTIAN) {
    const float r1 = rng.next1D();
    const float r2 = rng.next1D();
    wi = lambert_dir(n, r1, r2);
    weight = alb;
  } else if (bsdf == MUPSI_BSDF_NULL) {
    wi = neg(wo);
    weight = mk(1, 1, 1);
  } else {
    const float k = 2.0f * dot(wo, n);
    wi = mk(fmaf(k, n.x, -wo.x), fmaf(k, n.y, -wo.y), fmaf(k, n.z, -wo.z));
    weight = mk(1, 1, 1);
  }
}

// Guided Lambertian sample (one-sample MIS of the guide mixture m and the cosine
// lobe, with the vertex's learned selection probability alpha = m[40]):
// q = alpha p_guide + (1 - alpha) cos/pi, and the albedo-free weight
// f cos / q = (cos/pi) / q (0 below the surface).  u0 picks the technique; the
// guide branch takes u1..u3, the cosine branch takes (u1, u2) as its (r1, r2).
__device__ __forceinline__ void guided_eval(const float* m, F3 n, F3 wi, float& q, float& wfac) {
  const float alpha = m[guide::kAlphaIdx];
  const float cosv = dot(wi, n);
  const float pg = guide::mixture_pdf(m, wi.x, wi.y, wi.z);
  const float kInvPi = 0.318309886183790671538f;
  const float pb = fmaxf(cosv, 0.0f) * kInvPi;
  q = alpha * pg + (1.0f - alpha) * pb;
  wfac = (cosv > 0.0f && q > 0.0f) ? (cosv * kInvPi) / q : 0.0f;
}
__device__ __forceinline__ F3 guided_dir(const float* m, F3 n, float u0, float u1, float u2,
                                         float u3) {
  if (u0 < m[guide::kAlphaIdx]) {
    F3 w;
    guide::mixture_sample(m, u1, u2, u3, w.x, w.y, w.z);
    return w;
  }
  return lambert_dir(n, u1, u2);
}
__device__ __forceinline__ void load_mix(const float* src, float* m) {
  const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
  for (int r = 0; r < guide::kMix / 4; ++r) {
    const float4 v = s4[r];
    m[4 * r] = v.x;
    m[4 * r + 1] = v.y;
    m[4 * r + 2] = v.z;
    m[4 * r + 3] = v.w;
  }
}

// ====================================================================== device: primitives

// intersectUnitSphere (sphere.cpp:11-30)
__device__ __forceinline__ float unit_sphere_t(float nearT, float farT, F3 O, F3 D) {
  const float a = sq3(D);
  const float b = dot(O, D);
  const float c = sq3(O) - 1.0f;
  const float disc = fmaf(b, b, -(a * c));
  if (disc < 0.0f) return -1.0f;
  const float sd = sqrtf(disc);
  float t = (-b - sd) / a;
  if (t > nearT && t < farT) return t;
  t = (-b + sd) / a;
  if (t > nearT && t < farT) return t;
  return -1.0f;
}
__device__ __forceinline__ void sphere_local(const DevPrim& pr, const Ray& r, F3& O, F3& D) {
  O = r.o;
  D = r.d;
  if (!pr.xf.identity) {
    O = xf_to_local(pr.xf, r.o);
    D = mul(r.d, pr.xf.inv_s);
  }
}

// Scene::intersect (scene.cpp:25-39) + intersectInfo of the closest primitive.
// kTri false: the scene is known to hold spheres only (no BVH code is compiled).
template <bool kTri = true>
__device__ __forceinline__ int scene_intersect(const DevScene& S, Ray& ray, F3& P, F3& Ng, float& uu, float& vv) {
  float mint = 3.402823466e+38f;
  int hitp = -1;
  uint32_t hit_tri = 0;
  float hu = 0, hv = 0;
  for (int i = 0; i < S.n_prims; ++i) {
    const DevPrim& pr = S.prims[i];
    bool h = false;
    uint32_t ti = 0;
    float u = 0, v = 0;
    if (!kTri || pr.type == MUPSI_PRIM_SPHERE) {
      F3 O, D;
      sphere_local(pr, ray, O, D);
      const float t = unit_sphere_t(ray.nearT, ray.farT, O, D);
      if (t >= 0.0f) {
        ray.farT = t;
        h = true;
      }
    } else {
      h = mesh_intersect(S.meshes[pr.mesh], ray, ti, u, v);
    }
    if (h && ray.farT < mint) {
      mint = ray.farT;
      hitp = i;
      hit_tri = ti;
      hu = u;
      hv = v;
    }
  }
  if (hitp < 0) return -1;
  P = madd(ray.d, ray.farT, ray.o);
  const DevPrim& pr = S.prims[hitp];
  if (!kTri || pr.type == MUPSI_PRIM_SPHERE) {
    const F3 pl = pr.xf.identity ? P : xf_to_local(pr.xf, P);
    const F3 ln = normalized(pl);
    Ng = pr.xf.identity ? ln : normalized(mul(ln, pr.xf.inv_s));
    uu = 0.0f;
    vv = 0.0f;
  } else {
    const DevMesh& m = S.meshes[pr.mesh];
    F3 ln;
    if (!m.vnormals) {
      ln = ld3(m.fnormals, hit_tri);
    } else {  // Mesh::interpolateNormal (mesh.cpp:163-172)
      const int32_t* f = m.faces + 3 * hit_tri;
      const float w = 1.0f - hu - hv;
      const F3 n0 = ld3(m.vnormals, f[0]), n1 = ld3(m.vnormals, f[1]), n2 = ld3(m.vnormals, f[2]);
      const F3 n = mk(fmaf(hv, n2.x, fmaf(hu, n1.x, w * n0.x)),
                      fmaf(hv, n2.y, fmaf(hu, n1.y, w * n0.y)),
                      fmaf(hv, n2.z, fmaf(hu, n1.z, w * n0.z)));
      const float len = norm(n);
      ln = len > 1e-12f ? divs(n, len) : ld3(m.fnormals, hit_tri);
    }
    Ng = mesh_world_normal(m, ln);
    if (!m.texcoords || !m.face_tex) {
      uu = hu;
      vv = hv;
    } else {  // Mesh::interpolateTexcoord (mesh.cpp:174-181)
      const int32_t* t = m.face_tex + 3 * hit_tri;
      const float w = 1.0f - hu - hv;
      uu = fmaf(hv, m.texcoords[2 * t[2]], fmaf(hu, m.texcoords[2 * t[1]], w * m.texcoords[2 * t[0]]));
      vv = fmaf(hv, m.texcoords[2 * t[2] + 1],
                fmaf(hu, m.texcoords[2 * t[1] + 1], w * m.texcoords[2 * t[0] + 1]));
    }
  }
  return hitp;
}

// Scene::occluded (scene.cpp:41-48)
template <bool kTri = true>
__device__ __forceinline__ bool scene_occluded(const DevScene& S, const Ray& r) {
  for (int i = 0; i < S.n_prims; ++i) {
    const DevPrim& pr = S.prims[i];
    if (!kTri || pr.type == MUPSI_PRIM_SPHERE) {
      F3 O, D;
      sphere_local(pr, r, O, D);
      if (unit_sphere_t(r.nearT, r.farT, O, D) >= 0.0f) return true;
    } else if (mesh_occluded(S.meshes[pr.mesh], r)) {
      return true;
    }
  }
  return false;
}

struct LightSample {
  F3 w, d;
  float dist, pdf;
};

// Sphere::sampleDirect (sphere.cpp:91-133)
__device__ bool sphere_sample_direct(const DevPrim& pr, F3 pw, Pcg& rng, LightSample& ls) {
  const F3 pl = pr.xf.identity ? pw : xf_to_local(pr.xf, pw);
  F3 dir = neg(pl);
  const float dist2 = sq3(dir);
  const float d2 = dist2 - 1.0f;
  if (d2 < 1e-6f) return false;
  const float d = sqrtf(d2);
  const float dist = sqrtf(dist2);
  dir = divs(dir, dist);
  const float cosT = d / dist;
  const float x1 = rng.next1D();
  const float x2 = rng.next1D();
  const float phi = (float)((double)(x1 * 2) * M_PI);
  const float z = fmaf(x2, 1.0f - cosT, cosT);
  const float r = sqrtf(fmaxf(fmaf(-z, z, 1.0f), 0.0f));
  const F3 sd = mk((float)cos((double)phi) * r, (float)sin((double)phi) * r, z);
  const float B = dist * sd.z;
  const float det = sqrtf(fmaxf(fmaf(B, B, -d2), 0.0f));
  ls.dist = B - det;
  ls.d = to_global(make_frame(dir), sd);
  if (!pr.xf.identity) {
    const F3 spw = xf_to_world(pr.xf, madd(ls.d, ls.dist, pl));
    const F3 wd = sub(spw, pw);
    ls.dist = norm(wd);
    ls.d = divs(wd, ls.dist);
  }
  ls.pdf = (float)(1.0f / (2.0f * M_PI * (1.0f - cosT)));
  const F3 em = pr.has_em ? mk(pr.em[0], pr.em[1], pr.em[2]) : mk(0, 0, 0);
  ls.w = divs(em, fmaxf(ls.pdf, 1e-6f));
  return true;
}

// Mesh::sampleDirect (mesh.cpp:252-296): face by std::lower_bound on the area
// CDF, a point by sqrt(r1) barycentrics, pdf 1/total local area (area measure).
__device__ bool mesh_sample_direct(const DevPrim& pr, const DevMesh& m, F3 p, Pcg& rng,
                                   LightSample& ls) {
  if (m.n_faces <= 0 || !(m.total_area > 0.0f) || !m.face_cdf) return false;
  const float r = rng.next1D();
  int lo = 0, hi = m.n_faces;
  while (lo < hi) {  // first face with cdf >= r
    const int mid = (lo + hi) >> 1;
    if (m.face_cdf[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  const int fi = lo >= m.n_faces ? m.n_faces - 1 : lo;
  const float r1 = rng.next1D();
  const float r2 = rng.next1D();
  const float sr1 = sqrtf(r1);
  const float bu = 1.0f - sr1, bv = r2 * sr1;
  const int32_t* f = m.faces + 3 * fi;
  const F3 v0 = ld3(m.verts, f[0]), v1 = ld3(m.verts, f[1]), v2 = ld3(m.verts, f[2]);
  const F3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  const F3 pl = mk(fmaf(e2.x, bv, fmaf(e1.x, bu, v0.x)), fmaf(e2.y, bv, fmaf(e1.y, bu, v0.y)),
                   fmaf(e2.z, bv, fmaf(e1.z, bu, v0.z)));
  const F3 pw = m.xf.identity ? pl : xf_to_world(m.xf, pl);
  const F3 dir = sub(pw, p);
  const float dist = sqrtf(sq3(dir));
  ls.d = divs(dir, dist);
  ls.dist = dist;
  ls.pdf = 1.0f / m.total_area;
  const F3 em = pr.has_em ? mk(pr.em[0], pr.em[1], pr.em[2]) : mk(0, 0, 0);
  ls.w = divs(em, fmaxf(ls.pdf, 1e-6f));
  return true;
}

// Scene::chooseLight (scene.cpp:50-61): no draw when the scene has no emitter.
template <bool kTri = true>
__device__ __forceinline__ bool choose_light(const DevScene& S, F3 p, Pcg& rng, LightSample& ls) {
  if (S.n_prims == 0 || S.n_lights == 0) return false;
  const int idx = (int)(rng.nextI() % (uint32_t)S.n_lights);
  const DevPrim& pr = S.prims[S.lights[idx]];
  if (!kTri || pr.type == MUPSI_PRIM_SPHERE) return sphere_sample_direct(pr, p, rng, ls);
  return mesh_sample_direct(pr, S.meshes[pr.mesh], p, rng, ls);
}

// ====================================================================== device: march

// Certain-sign test of one grid point (see k_march).  |psi| <= gp.bound and the
// conditioning term is bounded by cond_bound, so wherever |mu| provably exceeds
// B = their sum plus a rounding margin the sign of f is the sign of mu.
//  - sphere mean: a shell test on |p - c|^2;
//  - mesh mean: mu = (inside ? -1 : 1) * dist and dist >= scale * dist(p_local,
//    root AABB).  Outside the +x shadow slab of the root box the parity ray
//    cannot reach the box, so inside() is exactly false; inside the slab one
//    inside() call (the reference's own parity test) decides the sign.
struct SkipInfo {
  int kind;            // 0 none, 1 sphere mean, 2 mesh mean
  float outer2, inner2;
  // sphere mean: every grid point that classify_point leaves uncertain has t in
  // [ev0, ev1] and outside (in0, in1); such t need no classification (an exact
  // evaluation takes the decision a certain sign would)
  float ev0, ev1, in0, in1;
  float Bm;            // mesh: B + margin (world units)
  F3 cb;               // mesh: centre of the world bounding ball of the root box
  float rb;            // mesh: its radius
  float wmin[3], wmax[3];  // mesh: root box in world space (s > 0)
};

// kMesh false: the scene's mean is known to be a sphere.
template <bool kMesh = true>
__device__ __forceinline__ SkipInfo make_skip(const DevScene& S, const Cond& c, F3 o, F3 d, float ft) {
  SkipInfo k;
  k.kind = 0;
  k.outer2 = INFINITY;
  k.inner2 = -1.0f;
  k.ev0 = INFINITY;
  k.ev1 = -INFINITY;
  k.in0 = INFINITY;
  k.in1 = -INFINITY;
  const float B = S.gp.bound + cond_bound(S.gp, c);
  if (!kMesh || S.mean_type == MUPSI_MEAN_SPHERE) {
    const F3 oc = sub(o, mk(S.mean_c[0], S.mean_c[1], S.mean_c[2]));
    const float m = 0.01f + 2e-6f * (norm(oc) + ft + S.mean_r + B);
    const float ro = S.mean_r + B + m;
    k.outer2 = ro * ro;
    const float ri = S.mean_r - B - m;
    k.inner2 = ri > 0.0f ? ri * ri : -1.0f;
    k.kind = 1;
    // The t-ranges of the ray inside the ball of radius ro (widened) and inside
    // that of radius ri (narrowed), in FP64.  A grid point's classification
    // compares the FP32 |p - c|^2 with outer2 / inner2; p = o + d t and p - c
    // carry errors below 1e-6 (|o| + |c| + |t|) per component and the square a
    // few ulps, far inside the widening, so every uncertain point lies in
    // [ev0, ev1] and outside (in0, in1).
    const double ox = oc.x, oy = oc.y, oz = oc.z, dx = d.x, dy = d.y, dz = d.z;
    const double a = dx * dx + dy * dy + dz * dz;
    const double b = ox * dx + oy * dy + oz * dz;
    const double cc = ox * ox + oy * oy + oz * oz;
    const double slack =
        1e-3 + 1e-5 * ((double)norm(o) + (double)norm(sub(o, oc)) + (double)ft + ro);
    if (a > 0.0) {
      const double rw = (double)ro * (1.0 + 1e-5) + slack;
      const double disc = b * b - a * (cc - rw * rw);
      // (the widenings w dwarf the rounding of a product with 1/a)
      const double ia = 1.0 / a;
      if (disc >= 0.0) {
        const double sq = sqrt(disc);
        const double w = 1e-5 * (1.0 + fabs(b) + sq) * ia + 1e-4;
        k.ev0 = (float)((-b - sq) * ia - w);
        k.ev1 = (float)((-b + sq) * ia + w);
      }
      if (ri > 0.0f) {
        const double rn = (double)ri * (1.0 - 1e-5) - slack;
        const double disci = b * b - a * (cc - rn * rn);
        if (rn > 0.0 && disci > 0.0) {
          const double sq = sqrt(disci);
          const double w = 1e-5 * (1.0 + fabs(b) + sq) * ia + 1e-4;
          const double i0 = (-b - sq) * ia + w, i1 = (-b + sq) * ia - w;
          if (i0 < i1) {
            k.in0 = (float)i0;
            k.in1 = (float)i1;
          }
        }
      }
    }
    return k;
  }
  if constexpr (!kMesh) return k;
  const DevMesh& M = S.mean_mesh;
  if (M.n_nodes == 0 || !(M.xf.s > 0.0f)) return k;
  for (int a = 0; a < 3; ++a) {
    k.wmin[a] = fmaf(M.xf.s, M.lbmin[a], M.xf.pos[a]);
    k.wmax[a] = fmaf(M.xf.s, M.lbmax[a], M.xf.pos[a]);
  }
  k.cb = mk(0.5f * (k.wmin[0] + k.wmax[0]), 0.5f * (k.wmin[1] + k.wmax[1]),
            0.5f * (k.wmin[2] + k.wmax[2]));
  k.rb = 0.5f * norm(mk(k.wmax[0] - k.wmin[0], k.wmax[1] - k.wmin[1], k.wmax[2] - k.wmin[2])) * 1.0001f;
  k.Bm = B + 0.01f + 2e-6f * (norm(sub(o, k.cb)) + ft + k.rb + B);
  k.kind = 2;
  return k;
}

// Returns +1 / -1 when the sign of f(p) is certain, 0 otherwise.  *final: no later
// grid point of the ray can have another certain-or-uncertain sign (the march can
// stop once this sign equals the origin's).
// kMesh false: the scene's mean is known not to be a mesh (no BVH code).
// ins: Mesh::inside(p) when the caller knows it (0 / 1), else -1.
template <bool kMesh = true>
__device__ __forceinline__ int classify_point(const DevScene& S, const SkipInfo& k, F3 p, F3 d,
                                              bool* final, int ins = -1) {
  *final = false;
  if (k.kind == 1) {
    const F3 r = sub(p, mk(S.mean_c[0], S.mean_c[1], S.mean_c[2]));
    const float q2 = sq3(r);
    if (q2 > k.outer2) {
      *final = dot(d, r) >= 0.0f;
      return 1;
    }
    return q2 < k.inner2 ? -1 : 0;
  }
  if (kMesh && k.kind == 2) {
    const DevMesh& M = S.mean_mesh;
    const F3 pl = xf_to_local(M.xf, p);
    const float dx = fmaxf(fmaxf(M.lbmin[0] - pl.x, 0.0f), pl.x - M.lbmax[0]);
    const float dy = fmaxf(fmaxf(M.lbmin[1] - pl.y, 0.0f), pl.y - M.lbmax[1]);
    const float dz = fmaxf(fmaxf(M.lbmin[2] - pl.z, 0.0f), pl.z - M.lbmax[2]);
    const float dw = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))) * M.xf.s;
    if (!(dw > k.Bm)) return 0;
    // the root-box test of the +x parity ray (AABB::intersect with d = (1, 0, 0))
    if (ins >= 0) {
      if (ins) return -1;
      if (mesh_in_parity_slab(M, pl)) return 1;
    } else if (mesh_in_parity_slab(M, pl)) {
      return mesh_inside(M, p) ? -1 : 1;
    }
    const F3 rc = sub(p, k.cb);
    const float e = 1e-3f * (1.0f + fabsf(p.x) + fabsf(p.y) + fabsf(p.z));
    const bool leaves = (p.x > k.wmax[0] + e && d.x >= 0.0f) || (p.y > k.wmax[1] + e && d.y >= 0.0f) ||
                        (p.y < k.wmin[1] - e && d.y <= 0.0f) || (p.z > k.wmax[2] + e && d.z >= 0.0f) ||
                        (p.z < k.wmin[2] - e && d.z <= 0.0f);
    *final = leaves && dot(d, rc) >= 0.0f && norm(rc) > k.rb + k.Bm;
    return 1;
  }
  return 0;
}

// The medium's field and gradient at p, any medium kind: GPMedium (mean + psi +
// conditioning), SDFMedium (the mean function alone, sdfmedium.cpp:10-14) or
// ExternalGPMedium (grids + hash noise; the conditioning state is ignored,
// externalgpmedium.cpp:26-29,60-63).
template <bool kGrad>
__device__ __forceinline__ float medium_field(const DevScene& S, F3 p, uint32_t seed, const Cond& c,
                                              F3* grad) {
  if (S.medium_kind == MUPSI_MEDIUM_SDF) return mean_eval<kGrad>(S, p, grad);
  if (S.medium_kind == MUPSI_MEDIUM_GRID) return grid_medium_eval<kGrad>(S, p, seed, grad);
  return field_eval<kGrad>(S, p, seed, c, grad);
}

// std::max(a, b) of the reference's step rules: (a < b) ? b : a
__device__ __forceinline__ float std_max(float a, float b) { return a < b ? b : a; }

// SDFMedium::sampleDistance / transmittance (sdfmedium.cpp:16-98): sphere
// tracing on the mean function, step max(|f|, 1e-3), at most 512 steps; a sign
// change (f_prev * f < 0) is refined by 16 bisection steps and the hit is hi.
// iters: steps taken.  Transmittance takes the same decision without the
// bisection.
template <bool kMesh = true>
__device__ __forceinline__ bool sdf_march(const DevScene& S, F3 o, F3 d, float nt, float ft, float& t_hit,
                          F3& p_hit, int& iters) {
  float t = nt;
  float fc = mean_eval<false, kMesh>(S, o, nullptr);
  int iter = 0;
  while (t < ft && iter < 512) {
    const F3 p = madd(d, t, o);
    const float fp = fc, tp = t;
    fc = mean_eval<false, kMesh>(S, p, nullptr);
    if (fp * fc < 0.0f) {
      float lo = tp, hi = t, flo = fp;
      for (int b = 0; b < 16; ++b) {
        const float mid = (lo + hi) * 0.5f;
        const float fm = mean_eval<false, kMesh>(S, madd(d, mid, o), nullptr);
        if (flo * fm <= 0.0f) {
          hi = mid;
        } else {
          lo = mid;
          flo = fm;
        }
      }
      t_hit = hi;
      p_hit = madd(d, hi, o);
      iters = iter + 1;
      return true;
    }
    t += std_max(fabsf(fc), 1e-3f);
    ++iter;
  }
  iters = iter;
  return false;
}

// ExternalGPMedium::sampleDistance / transmittance (externalgpmedium.cpp:68-130):
// the same stepping with min step 0.005 and at most 1024 steps, GPMedium's sign
// test (sign(x) = x > 0 and |f_prev - f| > 1e-6), bisection from
// t - max(|f_prev|, 0.005) to t keeping f_prev's sign at lo.
__device__ __forceinline__ bool grid_march(const DevScene& S, F3 o, F3 d, float nt, float ft, uint32_t seed,
                           float& t_hit, F3& p_hit, int& iters) {
  const float min_step = 0.005f;
  float t = nt;
  float fprev = grid_medium_eval<false>(S, o, seed, nullptr);
  int iter = 0;
  for (; iter < 1024 && t < ft; ++iter) {
    const F3 p = madd(d, t, o);
    const float fc = grid_medium_eval<false>(S, p, seed, nullptr);
    if ((fc > 0.0f) != (fprev > 0.0f) && fabsf(fprev - fc) > 1e-6f) {
      float lo = t - std_max(fabsf(fprev), min_step);
      float hi = t;
      for (int b = 0; b < 16; ++b) {
        const float mid = (lo + hi) * 0.5f;
        const float fm = grid_medium_eval<false>(S, madd(d, mid, o), seed, nullptr);
        if ((fm > 0.0f) == (fprev > 0.0f)) lo = mid;
        else hi = mid;
      }
      t_hit = hi;
      p_hit = madd(d, hi, o);
      iters = iter + 1;
      return true;
    }
    fprev = fc;
    t += std_max(fabsf(fc), min_step);
  }
  iters = iter;
  return false;
}

// The march of a medium other than the GP medium (hit flag, t, point).  Every
// function taking the scene by reference is inlined: an out-of-line call needs
// the scene's address, and the kernel would copy its scene argument to scratch.
// kMesh false: the mean is known to be a sphere.
template <bool kMesh = true>
__device__ __forceinline__ bool alt_march(const DevScene& S, F3 o, F3 d, float nt, float ft,
                                       uint32_t seed, float& t_hit, F3& p_hit, int& iters) {
  if (S.medium_kind == MUPSI_MEDIUM_SDF) return sdf_march<kMesh>(S, o, d, nt, ft, t_hit, p_hit, iters);
  return grid_march(S, o, d, nt, ft, seed, t_hit, p_hit, iters);
}
// Gradient of a medium other than the GP medium at p.
template <bool kMesh = true>
__device__ __forceinline__ F3 alt_grad(const DevScene& S, F3 p, uint32_t seed) {
  F3 g = mk(0, 0, 0);
  if (S.medium_kind == MUPSI_MEDIUM_SDF) (void)mean_eval<true, kMesh>(S, p, &g);
  else (void)grid_medium_eval<true>(S, p, seed, &g);
  return g;
}

// Serial march of one ray (used for NEE shadow rays in k_shade): the same
// decisions as k_march without the certain-sign skip.
template <bool kMesh = true>
__device__ __forceinline__ bool march_serial(const DevScene& S, const Ray& ray, uint32_t seed,
                                             const Cond& c) {
  if (S.medium_kind != MUPSI_MEDIUM_GP) {
    float th;
    F3 ph;
    int it;
    return alt_march<kMesh>(S, ray.o, ray.d, ray.nearT, fminf(ray.farT, 2000.0f), seed, th, ph, it);
  }
  const float nt = ray.nearT, ft = fminf(ray.farT, 2000.0f);
  const float dt = fminf(0.5f, (ft - nt) / 100);
  float t = nt;
  const float fprev = field_eval<false, kMesh>(S, ray.o, seed, c, nullptr);
  const bool sp = fprev > 0.0f;
  while (t < ft) {
    const float fc = field_eval<false, kMesh>(S, madd(ray.d, t, ray.o), seed, c, nullptr);
    if ((fc > 0.0f) != sp && fabsf(fprev - fc) > 1e-6f) return true;
    if (!(t + dt > t)) return false;
    t += dt;
  }
  return false;
}

struct MarchArgs {
  const int* queue;  // null: identity
  const int* count_dev;
  int count_host;
  const float *ox, *oy, *oz, *dx, *dy, *dz, *nt, *ft;  // ft null: FLT_MAX
  const uint32_t* seed;
  const uint32_t* flags;  // render: conditioning-active bit
  const uint8_t* active;  // batch API: conditioning active, or null
  const float *cx, *cy, *cz, *cu, *cgx, *cgy, *cgz;
  uint8_t* hit;
  float *t, *px, *py, *pz;
  int32_t* steps;  // t-grid points visited up to and including the hit, or null
  int* work;
  unsigned long long* stats;  // [0] field evals, [1] t-grid points (reference work), [2] rays
  const DevScene* scenes;     // batched launch: per-scene constants, indexed by sid[ray]
  const uint16_t* sid;
  // fuse != 0: the marcher shades the rays' GP-medium hits itself and marches on
  // (render wavefront of a sphere-mean GP medium without emitters or guide);
  // the path state it updates, and path id -> (x, y, k) as in PathSrc
  int fuse;
  uint64_t* rng;
  float *Tx, *Ty, *Tz;
  int src_W, src_npix, src_k_begin;
  const int* src_k_begin_dev;
  const int32_t *src_xs, *src_ys, *src_ks;
};
static_assert(std::is_standard_layout<MarchArgs>::value, "MarchArgs fields are read by offset");

// A field of a kernel's FIRST argument, read from the kernarg segment where it is
// used.  Pointers that a persistent kernel needs only when it takes or retires a
// ray are otherwise held in scalar registers for the whole kernel, and the
// scalar registers they take are spilled to vector-register lanes that the hot
// loop then reads back with VALU instructions.
template <typename Args, typename T, size_t kOff>
__device__ __forceinline__ T kernarg_field() {
  using CP = const char __attribute__((address_space(4)))*;
  const CP base = (CP)__builtin_amdgcn_kernarg_segment_ptr();
  return *(const volatile T __attribute__((address_space(4)))*)(base + kOff);
}
// MARCH_ARG is valid only in kernels whose first parameter is the MarchArgs
// (kernel arguments are laid out from offset 0 in declaration order).
#define MARCH_ARG(field) kernarg_field<MarchArgs, decltype(MarchArgs::field), offsetof(MarchArgs, field)>()

// GPMedium::sampleDistance / transmittance (gpmedium.cpp:46-129), lane per ray.
//
// Reference semantics: dt = min(0.5, (min(farT,2000) - nearT)/100); f_prev =
// f(origin) stays fixed; t starts at nearT and advances by FP32 t += dt; the
// first grid point with sign(f) != sign(f_prev) (sign(x) = x > 0) and
// |f_prev - f| > 1e-6 is the hit; the lerp refinement result is discarded.
//
// Certain-sign skip (same outputs, fewer evaluations): |psi| <= gp.bound and
// |conditioning| <= cond_bound everywhere, so for a sphere mean every grid point
// with | |p-c| - R | > B + margin has a known sign.  Such points are not
// evaluated: same sign as f_prev -> no hit; opposite sign -> the hit is there.
// Past the outer shell and moving away, no later point can change sign.
template <bool kMulti>
__global__ __launch_bounds__(256) void k_march(DevScene S0, MarchArgs a) {
  const int n = a.count_dev ? *a.count_dev : a.count_host;
  const int lane = __lane_id();
  int phase = 0;  // 0 needs a ray, 1 evaluate origin, 2 stepping, 3 no work left
  int rid = 0;
  F3 o = mk(0, 0, 0), d = mk(0, 0, 0);
  float t = 0.f, ft = 0.f, dt = 0.f, nt = 0.f, fprev = 0.f;
  bool sp = false;
  uint32_t seed = 0;
  Cond c;
  c.active = false;
  SkipInfo sk;
  sk.kind = 0;
  int nsteps = 0;
  unsigned long long evals = 0, refsteps = 0, rays = 0;
  int sidx = 0;  // batched launch: scene of this lane's ray
  auto scene = [&]() -> const DevScene& {
    if constexpr (kMulti) return a.scenes[sidx];
    else return S0;
  };

  auto finish = [&](bool hit, F3 p) {
    a.hit[rid] = hit ? 1 : 0;
    if (hit) {
      a.t[rid] = t;
      a.px[rid] = p.x;
      a.py[rid] = p.y;
      a.pz[rid] = p.z;
      refsteps += nsteps;
    } else {
      refsteps += nsteps + (unsigned long long)fmaxf(0.0f, ceilf((ft - t) / dt));
    }
    if (a.steps) a.steps[rid] = hit ? nsteps : -1;
    phase = 0;
  };

  while (true) {
    bool do_eval = false;
    F3 q = mk(0, 0, 0);
    for (int round = 0; round < 2; ++round) {
      const bool need = phase == 0;
      const uint64_t m = __ballot(need);
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(a.work, __popcll(m));
        base = __shfl(base, leader);
        if (need) {
          const int qi = base + __popcll(m & ((1ull << lane) - 1ull));
          if (qi < n) {
            rid = a.queue ? a.queue[qi] : qi;
            if constexpr (kMulti) sidx = a.sid[rid];
            o = mk(a.ox[rid], a.oy[rid], a.oz[rid]);
            d = mk(a.dx[rid], a.dy[rid], a.dz[rid]);
            nt = a.nt[rid];
            ft = fminf(a.ft ? a.ft[rid] : 3.402823466e+38f, 2000.0f);
            dt = fminf(0.5f, (ft - nt) / 100);
            seed = a.seed[rid];
            c.active = a.flags ? (a.flags[rid] & kFlagCond) != 0 : (a.active && a.active[rid]);
            if (c.active) {
              c.C = mk(a.cx[rid], a.cy[rid], a.cz[rid]);
              c.u = a.cu[rid];
              c.g = mk(a.cgx[rid], a.cgy[rid], a.cgz[rid]);
            }
            ++rays;
            if (scene().medium_kind != MUPSI_MEDIUM_GP) {
              // SDF / grid media: the whole sphere-traced march of this ray here
              float th = 0.0f;
              F3 ph = mk(0, 0, 0);
              int it = 0;
              const bool hit = alt_march(scene(), o, d, nt, ft, seed, th, ph, it);
              a.hit[rid] = hit ? 1 : 0;
              if (hit) {
                a.t[rid] = th;
                a.px[rid] = ph.x;
                a.py[rid] = ph.y;
                a.pz[rid] = ph.z;
              }
              if (a.steps) a.steps[rid] = hit ? it : -1;
              evals += it + 1;
              refsteps += it;
              phase = 0;
            } else {
              sk = make_skip(scene(), c, o, d, ft);
              nsteps = 0;
              phase = 1;
            }
          } else {
            phase = 3;
          }
        }
      }
      if (phase == 1) {
        q = o;
        do_eval = true;
      } else if (phase == 2 && !do_eval) {
        while (true) {
          if (!(t < ft)) {
            finish(false, q);
            break;
          }
          const F3 p = madd(d, t, o);
          bool fin;
          const int cs = classify_point(scene(), sk, p, d, &fin);
          if (cs == 0) {
            q = p;
            do_eval = true;
            break;
          }
          ++nsteps;
          if ((cs > 0) != sp) {
            finish(true, p);
            break;
          }
          if (fin || !(t + dt > t)) {  // no later sign change / the grid cannot advance
            finish(false, p);
            break;
          }
          t += dt;
        }
      }
      if (__ballot(phase == 0) == 0) break;
    }
    if (__ballot(phase != 3) == 0) break;
    if (do_eval) {
      const float f = field_eval<false>(scene(), q, seed, c, nullptr);
      ++evals;
      if (phase == 1) {
        fprev = f;
        sp = f > 0.0f;
        t = nt;
        phase = 2;
      } else {
        ++nsteps;
        if ((f > 0.0f) != sp && fabsf(fprev - f) > 1e-6f) {
          finish(true, q);
        } else if (!(t + dt > t)) {  // the grid cannot advance: no further point exists
          finish(false, q);
        } else {
          t += dt;
        }
      }
    }
  }
  evals = wave_sum(evals);
  refsteps = wave_sum(refsteps);
  rays = wave_sum(rays);
  if (lane == 0 && a.stats) {
    atomicAdd(&a.stats[0], evals);
    atomicAdd(&a.stats[1], refsteps);
    atomicAdd(&a.stats[2], rays);
  }
}

// f at q for the window marchers, given psi_t = psi_scale * (the window's impulse
// sum).  The product replaces the reference's sigma * sum / sqrt(Var), the
// hardware exp2 and sqrt the correctly rounded expf and sqrtf of the
// conditioning kernel and the sphere mean: differences of an ulp or two, which
// no unambiguous decision sees (the psi sum's order differs anyway).  The mean is
// evaluated exactly unless, for a mesh mean, bounds of the distance decide the
// sign: with mu = s d, d in [lo, hi] and s from inside() (called only where the
// parity ray can reach the mesh), f lies in [s lo, s hi] + (psi term + conditioning
// term) up to rounding; when that interval keeps clear of 0 by a margin, the
// returned value carries the sign with a magnitude no larger than |f|.  Every
// decision the marcher takes from it (sign, and |f_prev - f| > 1e-6 across a
// sign change) is the one the exact value gives.
template <bool kMesh = true>
__device__ __forceinline__ float cached_field(const DevScene& S, const DevGP& G, F3 q, float psi_t,
                                              const Cond& c, float psi_k, int ins = -1) {
  float cond_t = 0.0f;
  if (c.active) {  // h(C, q) = exp2(|q - C|^2 psi_k), the impulses' hardware exp2
    const F3 r = sub(q, c.C);
    const float h = __builtin_amdgcn_exp2f(sq3(r) * psi_k);
    cond_t = fmaf(c.u, h, dot(c.g, mul(r, h * G.nk2)));
  }
  if (kMesh && S.mean_type == MUPSI_MEAN_MESH && S.mean_mesh.xf.s > 0.0f) {
    const DevMesh& M = S.mean_mesh;
    const F3 pl = xf_to_local(M.xf, q);
    float lo, hi;
    mesh_dist_bounds_local(M, pl, lo, hi);
    const float hi_local = hi;
    lo = lo * M.xf.s * (1.0f - 1e-5f);
    hi = hi * M.xf.s * (1.0f + 1e-5f);
    const float e = psi_t + cond_t;
    const float margin = 1e-3f + 1e-5f * (fabsf(psi_t) + fabsf(cond_t) + lo);
    const auto inside = [&]() {
      return ins >= 0 ? ins != 0 : (mesh_in_parity_slab(M, pl) && mesh_inside(M, q));
    };
    if (lo > fabsf(e) + margin) {  // |mu| alone decides: no sign query beyond the parity ray
      const float sg = inside() ? -1.0f : 1.0f;
      return sg * (lo - fabsf(e));
    }
    if (hi < INFINITY) {
      const float sg = inside() ? -1.0f : 1.0f;
      const float a = sg * lo + e, b = sg * hi + e;   // f lies between a and b
      const float fmin = fminf(a, b), fmax = fmaxf(a, b);
      if (fmin > margin) return fmin;
      if (fmax < -margin) return fmax;
      // within the interval a zero crossing is possible: evaluate mu exactly
      uint32_t ti = 0;
      F3 o;
      float f = sg * mesh_dist(M, q, ti, o, NoCount(), hi_local) + psi_t;  // mu = sg * d
      if (c.active) f = f + cond_t;
      return f;
    }
  }
  // the sphere mean of mean_eval, without the mesh code where no mesh can occur
  const float mu =
      kMesh ? mean_eval<false>(S, q, nullptr)
            : __builtin_amdgcn_sqrtf(sq3(sub(q, mk(S.mean_c[0], S.mean_c[1], S.mean_c[2])))) - S.mean_r;
  float f = mu + psi_t;
  if (c.active) f = f + cond_t;
  return f;
}

struct CachedConsts {
  float inv_kr;     // 1 / kr (rounded)
  int kr_pow2;      // kr is a power of two: q * inv_kr == q / kr exactly
  float psi_k;      // -log2(e) / L^2: h = exp2(|p - s|^2 * psi_k)
  float psi_scale;  // sigma / sqrt(Var): psi = psi_scale * (impulse sum)
};

// Sum over the GS lanes of a group by an xor butterfly (16, 8, 4, 2, 1 for 32
// lanes), every lane ending with the same total: ds_swizzle in bit mode
// (and 0x1f, xor k) exchanges within 32-lane halves without address arithmetic.
// Smaller groups start the butterfly at xor GS / 2.
template <int GS = 32>
__device__ __forceinline__ float group_sum(float v) {
  if constexpr (GS >= 32)
    v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
  if constexpr (GS >= 16)
    v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x201F));
  v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x101F));
  v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x081F));
  v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x041F));
  return v;
}

// The xor-k neighbour's value within 32-lane groups (ds_swizzle, bit mode).
template <int K>
__device__ __forceinline__ float swizzle_xor(float v) {
  static_assert(K == 1 || K == 2 || K == 4 || K == 8 || K == 16, "xor within 32 lanes");
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (K << 10) | 0x1F));
}

// Two sums over a group at once: lanes gl < GS/2 start with (a = their part of
// sum 1, b = their part of sum 2), the others with (their part of 2, of 1).  The
// first step exchanges across the halves, the rest is group_sum's butterfly
// within each half: lanes gl < GS/2 end with sum 1, the others with sum 2, each
// added in group_sum's order up to commutation.
template <int GS>
__device__ __forceinline__ float group_sum2(float a, float b) {
  float v = a + swizzle_xor<GS / 2>(b);
  if constexpr (GS >= 32) v += swizzle_xor<8>(v);
  if constexpr (GS >= 16) v += swizzle_xor<4>(v);
  if constexpr (GS >= 8) v += swizzle_xor<2>(v);
  v += swizzle_xor<1>(v);
  return v;
}

// This lane's part of the window sum: slots gl, gl + GS, ... of the 27 x PPC.
template <int PPC, int GS = kGroup>
__device__ __forceinline__ float window_sum(const float4* W, F3 q, int gl, float psi_k) {
  float acc = 0.0f;
#pragma unroll
  for (int sl = gl; sl < 27 * PPC; sl += GS) {
    const float4 im = W[sl];
    const float ex = q.x - im.x, ey = q.y - im.y, ez = q.z - im.z;
    const float r2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
    acc = fmaf(im.w, __builtin_amdgcn_exp2f(r2 * psi_k), acc);
  }
  return acc;
}

// This lane's parts of the window sums at q and q2 (same slots, same order);
// without a pair anywhere in the wave the second sum is not formed.
template <int PPC, int GS>
__device__ __forceinline__ void window_sums(const float4* W, F3 q, F3 q2, bool two, int gl,
                                            float psi_k, float& acc, float& acc2) {
  if (!two) {
    acc = window_sum<PPC, GS>(W, q, gl, psi_k);
    acc2 = acc;
    return;
  }
  acc = 0.0f;
  acc2 = 0.0f;
#pragma unroll
  for (int sl = gl; sl < 27 * PPC; sl += GS) {
    const float4 im = W[sl];
    const float ex = q.x - im.x, ey = q.y - im.y, ez = q.z - im.z;
    const float r2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
    acc = fmaf(im.w, __builtin_amdgcn_exp2f(r2 * psi_k), acc);
    const float fx = q2.x - im.x, fy = q2.y - im.y, fz = q2.z - im.z;
    const float s2 = fmaf(fx, fx, fmaf(fy, fy, fz * fz));
    acc2 = fmaf(im.w, __builtin_amdgcn_exp2f(s2 * psi_k), acc2);
  }
}

// Sphere mean: t lies where no classification can decide the sign.
__device__ __forceinline__ bool uncertain_t(const SkipInfo& k, float t) {
  return t >= k.ev0 && t <= k.ev1 && !(t > k.in0 && t < k.in1);
}

// The grid point after t when the march reaches it and it must be evaluated too
// (sphere means); q2 receives it.
__device__ __forceinline__ bool next_uncertain(const DevScene& S, const SkipInfo& k, F3 d, F3 o,
                                               float t, float dt, float ft, F3& q2) {
  const float t2 = t + dt;
  if (!(t2 > t) || !(t2 < ft) || k.kind != 1) return false;
  q2 = madd(d, t2, o);
  if (uncertain_t(k, t2)) return true;
  bool fin;
  return classify_point<false>(S, k, q2, d, &fin) == 0;
}

// ====================================================================== device: Renewal+

// GPMedium::sampleCondition (gpmedium.cpp:133-162) as a 4x4 FP64 solve of
//   [[h(C,C), grad h(C,C)^T], [grad h(C,C), H_h(C,C)]] [u; g] =
//   [-mu(C)/sigma - psi(C); n - grad mu(C) - grad psi(C)]
// with psi the next realization.  At r = 0 the matrix is diag(1, -2/L^2 I),
// which gives the reference's closed form u = -mu/sigma - psi,
// g = (-L^2/2)(n - grad mu - grad psi).
// The FP64 system at C for the mean (mu, grad mu) and the next realization's
// psi (pv, grad psi) at C.  The kernel is evaluated at r = C - C = 0, where the
// off-diagonal entries h (-2/L^2) r_j and the r_j r_k parts of the Hessian are
// exactly zero: Gaussian elimination (partial pivoting keeps every pivot on
// the diagonal) leaves the right-hand side as it is and divides it by the
// diagonal, which is what this computes, in FP64, from the same entries.
__device__ __forceinline__ void renewal_solve_at(const DevGP& gp, F3 n, float mu, F3 mg, float pv,
                                                 F3 pg, double& u, double g[3]) {
  const double L2 = gp.L2;
  const double h = exp(-0.0 / L2);               // h(C, C)
  const double hjj = h * ((-2.0 / L2) + 0.0);    // H_h(C, C)_jj; r_j r_k = 0
  const double b0 = -(double)mu / (double)gp.sigma - (double)pv;
  u = b0 / h;
  g[0] = ((double)n.x - (double)mg.x - (double)pg.x) / hjj;
  g[1] = ((double)n.y - (double)mg.y - (double)pg.y) / hjj;
  g[2] = ((double)n.z - (double)mg.z - (double)pg.z) / hjj;
}

__device__ __forceinline__ void renewal_solve(const DevScene& S, F3 C, F3 n, uint32_t seed_next,
                                              double& u, double g[3],
                                              const MeanAt* mean = nullptr) {
  F3 pg;
  float pv;
  psi_eval<true>(S.gp, C, seed_next, pv, pg, nullptr);
  F3 mg;
  float mu;
  if (mean) {  // the same mean_eval<true>(S, C) the caller made
    mu = mean->mu;
    mg = mean->grad;
  } else {
    mu = mean_eval<true>(S, C, &mg);
  }
  renewal_solve_at(S.gp, n, mu, mg, pv, pg, u, g);
}

// A time before which every grid point o + d t lies in cell (wx, wy, wz) of size
// kr, whatever the rounding.  Cell index w along an axis (trunc toward zero)
// covers [w kr, (w+1) kr) for w > 0, ((w-1) kr, w kr] for w < 0 and (-kr, kr) for
// w = 0; the ray leaves through the face it moves toward.  The exit time is less
// a margin for the rounding of the face, of the exit time (an approximate
// reciprocal) and of the points' coordinates (relative 1e-5, far above each).
__device__ __forceinline__ float cell_exit_t(F3 o, F3 d, int wx, int wy, int wz, float kr) {
  const auto axis = [kr](float oa, float da, int w) {
    if (!(da != 0.0f)) return INFINITY;
    const int face = da > 0.0f ? (w >= 0 ? w + 1 : w) : (w > 0 ? w : w - 1);
    const float b = (float)face * kr;
    const float id = __builtin_amdgcn_rcpf(da);
    const float tb = (b - oa) * id;
    const float m = 1e-5f * ((fabsf(b) + fabsf(oa) + kr) * fabsf(id) + fabsf(tb));
    return tb - m;
  };
  return fminf(axis(o.x, d.x, wx), fminf(axis(o.y, d.y, wy), axis(o.z, d.z, wz)));
}

// A bound on |f(p) - f(q)| / |p - q| for p, q in one cell (the same 27-cell
// window): 1 for the sphere mean's distance, psi_scale x (27 ppc impulses) x
// max |grad h| for the window sum, and for the conditioning term u h + g . grad h
// |u| max|grad h| + |g| max|Hessian h| with max|Hessian h| <= 1.2132 |nk2|
// (|nk2| e^{-x} (1 + 2x), x = r^2 / L^2); widened by 0.1% for rounding.
__device__ __forceinline__ float skip_lipschitz(const DevGP& g, float psi_scale, int ppc,
                                                const Cond& c) {
  float l = 1.0f + fabsf(psi_scale) * (27.0f * (float)ppc) * g.grad_h_max;
  if (c.active) l += fabsf(c.u) * g.grad_h_max + 1.2132f * fabsf(g.nk2) * norm(c.g);
  return l * 1.001f;
}

// The window marcher: GPMedium::sampleDistance / transmittance
// (gpmedium.cpp:46-129) with one group of GS lanes per ray (64 / GS rays per
// wave).  The 27 x ppc impulses of the 3x3x3 cell window around the evaluation
// point live in LDS as a toroidal window indexed by cell mod 3; when the march
// point enters a new cell only the 9/3/1 newly entered cells are regenerated
// (one PCG jump-ahead per impulse, the wave's entered cells as one pool over its
// lanes).  Each lane sums its slots; a butterfly finishes the sum.  Per pass a
// group evaluates its next uncertain grid point and, for sphere means, the one
// after it when that lies in the same cell.
//   PPC      points per cell (kMulti: the LDS capacity, 8; each ray uses its own)
//   kMesh    mean meshes possible: inside() at the next GS grid points of a ray
//            is evaluated at once, one point per lane, and kept as a bit mask
//   kMulti   batched launch: every ray takes its scene's constants by sid[ray]
// The MarchArgs stay the first parameter: MARCH_ARG reads them at offset 0.
template <int PPC, bool kMesh, bool kMulti, int GS = kGroup>
__device__ __forceinline__ void march_window(const MarchArgs& a, const DevScene& S0,
                                             const CachedConsts& cc0, const CachedConsts* ccs) {
  static_assert(GS == 8 || GS == 16 || GS == 32, "a ray's group is 8, 16 or 32 lanes");
  constexpr uint64_t kGroupMask = (1ull << GS) - 1ull;
  constexpr int NG = 64 / GS;  // rays (groups) per wave
  __shared__ float4 win[NG][27 * PPC];
  // a group's window regeneration for the wave's pooled pass: cell (x, y, z), seed;
  // entered-cell pattern | window residues | points per cell; cell size
  __shared__ int4 regen_cell[NG];
  __shared__ int2 regen_aux[NG];
  // The regeneration tables, copied from constant memory once per block where
  // the window is small: a global load in every regenerating pass put two memory
  // round trips on the latency of a pass.  With the larger windows (mesh means,
  // batched launches) the extra LDS would cost resident blocks instead.
  constexpr bool kLdsTables = !kMesh && !kMulti && sizeof(win) <= 8192;
  __shared__ uint8_t new_off[kLdsTables ? 28 * 27 : 1];
  __shared__ ulonglong2 states[kLdsTables ? 4 * PPC : 1];
  const int n = a.count_dev ? *a.count_dev : a.count_host;
  // an empty queue (later bounce iterations where no path went back into the
  // medium): the block ends before any setup (the whole block, one wave, returns)
  if (n <= 0) return;
  if constexpr (kLdsTables) {
    for (int i = threadIdx.x; i < 28 * 27; i += 64) new_off[i] = kNewCells.off[i / 27][i % 27];
    if (threadIdx.x < 4 * PPC)
      states[threadIdx.x] = make_ulonglong2(kStates.a[threadIdx.x], kStates.c[threadIdx.x]);
    // (the first pass's barrier orders these writes before any read)
  }
  const int lane = threadIdx.x;
  const int grp = lane / GS;
  const int gl = lane & (GS - 1);
  float4* W = win[grp];
  // the group's scene (batched launches: that of its current ray)
  int sidx = 0;
  DevGP G = S0.gp;
  CachedConsts cc = cc0;
  int ppc = PPC;
  const auto scene = [&]() -> const DevScene& {
    if constexpr (kMulti) return a.scenes[sidx];
    else return S0;
  };
  // ray state: identical in the GS lanes of a group
  // 0 needs a ray, 1 evaluate origin, 2 stepping, 3 no work left, 4/5 the ray ended
  // (miss/hit) in this pass, 7 Renewal+ conditioning of a shaded hit
  int phase = 0;
  int rid = 0;
  F3 o = mk(0, 0, 0), d = mk(0, 0, 0);
  float t = 0.f, ft = 0.f, dt = 0.f, nt = 0.f, fprev = 0.f;
  SkipInfo sk;
  sk.kind = 0;
  bool sp = false;
  uint32_t seed = 0;
  Cond c;
  c.active = false;
  int nsteps = 0;
  // Lipschitz bound of f along the ray within one cell's window, and its
  // reciprocal made slightly smaller (reach below stays an underestimate)
  float lip = 0.f, inv_lip = 0.f;
  int wx = 0, wy = 0, wz = 0;
  bool wvalid = false;
  // mesh means whose cells span several grid steps: |psi| <= wbound at every
  // point of the window's cell (from the window's impulses; INFINITY: unknown),
  // and |conditioning term| <= condb everywhere on the ray
  float wbound = INFINITY, condb = 0.0f;
  uint32_t evals = 0, refsteps = 0, rays = 0;  // per lane; a launch stays far below 2^32
  // Medium hits shaded in the marcher (MarchArgs::fuse): a hit (phase 5) is
  // shaded at the end of its pass when the group's window is that of the hit
  // point's cell (the hit was evaluated); a hit that certain-sign classification
  // decided at a point that was not evaluated first takes a pass that moves the
  // window there (phase 6, same realization).  Phase 7 conditions the next ray
  // at C = c.C with normal c.g (the window is regenerated there with the next
  // realization in that pass), and next_ray marks a phase-0 group whose next
  // ray is already in its registers.
  constexpr bool kFuse = !kMesh && !kMulti;
  const bool fuse = kFuse && MARCH_ARG(fuse) != 0;
  bool next_ray = false;
  // mesh means: Mesh::inside at grid points gi0 .. gi0 + GS - 1 of the current ray
  int gi = 0, gi0 = -1;
  uint32_t imask = 0;
  const auto inside_here = [&]() -> int {
    if (gi0 < 0 || gi - gi0 >= GS) {
      float tl = t;  // the march's own f32 accumulation of t, gl steps ahead
      for (int i = 0; i < gl; ++i) tl += dt;
      const F3 pp = madd(d, tl, o);
      const DevMesh& M = scene().mean_mesh;
      const bool in = mesh_in_parity_slab(M, xf_to_local(M.xf, pp)) && mesh_inside(M, pp);
      imask = (uint32_t)((__ballot(in) >> (grp * GS)) & kGroupMask);
      gi0 = gi;
    }
    return (int)((imask >> (gi - gi0)) & 1u);
  };

  // A ray ends in this pass (phase 4: miss, 5: hit at the current t); its outputs
  // are written once, at the end of the pass.
  auto finish = [&](bool hit, F3) { phase = hit ? 5 : 4; };

  while (true) {
    bool do_eval = false;
    F3 q = mk(0, 0, 0);
    int qins = -1;  // Mesh::inside(q) if known
    // sphere means: the grid point after q, evaluated in the same pass when it is
    // uncertain too and lies in q's cell (same window)
    bool pair = false;
    F3 q2 = mk(0, 0, 0);
    // One fetch per pass: a group whose ray ended in this pass's stepping takes its
    // next ray in the next pass (a second fetch round cost more than the pass it
    // saves such a group).
    {
      const bool need = phase == 0 && !next_ray;
      const uint64_t m = __ballot(need && gl == 0);
      // a path whose hit the marcher shaded marches its next ray (set up in
      // registers) without going back to the queue
      bool take = phase == 0 && next_ray;
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(MARCH_ARG(work), __popcll(m));
        base = __shfl(base, leader);
        if (need) {
          const int qi = base + __popcll(m & ((1ull << (grp * GS)) - 1ull));
          if (qi < n) {
            const int* const queue = MARCH_ARG(queue);
            rid = queue ? queue[qi] : qi;
            if constexpr (kMulti) {
              sidx = MARCH_ARG(sid)[rid];
              G = a.scenes[sidx].gp;
              ppc = G.ppc;
              cc = ccs[sidx];
            }
            o = mk(MARCH_ARG(ox)[rid], MARCH_ARG(oy)[rid], MARCH_ARG(oz)[rid]);
            d = mk(MARCH_ARG(dx)[rid], MARCH_ARG(dy)[rid], MARCH_ARG(dz)[rid]);
            nt = MARCH_ARG(nt)[rid];
            const float* const ftp = MARCH_ARG(ft);
            ft = fminf(ftp ? ftp[rid] : 3.402823466e+38f, 2000.0f);
            seed = MARCH_ARG(seed)[rid];
            const uint32_t* const flags = MARCH_ARG(flags);
            const uint8_t* const act = MARCH_ARG(active);
            c.active = flags ? (flags[rid] & kFlagCond) != 0 : (act && act[rid]);
            if (c.active) {
              c.C = mk(MARCH_ARG(cx)[rid], MARCH_ARG(cy)[rid], MARCH_ARG(cz)[rid]);
              c.u = MARCH_ARG(cu)[rid];
              c.g = mk(MARCH_ARG(cgx)[rid], MARCH_ARG(cgy)[rid], MARCH_ARG(cgz)[rid]);
            }
            take = true;
          } else {
            phase = 3;
          }
        }
      }
      if (take) {
        next_ray = false;
        dt = fminf(0.5f, (ft - nt) / 100);
        sk = make_skip<kMesh>(scene(), c, o, d, ft);
        lip = skip_lipschitz(G, cc.psi_scale, ppc, c);
        inv_lip = (1.0f / lip) * 0.999999f;
        if constexpr (kMesh) condb = cond_bound(G, c);
        nsteps = 0;
        wvalid = false;  // new realization seed
        wbound = INFINITY;
        gi0 = -1;
        if (gl == 0) ++rays;
        phase = 1;
      }
      if (phase == 1) {
        q = o;
        do_eval = true;
      } else if (phase == 2 && !do_eval && t < ft && sk.kind == 1 && uncertain_t(sk, t)) {
        // the common case, outside the stepping loop: a point no classification decides
        q = madd(d, t, o);
        do_eval = true;
        if constexpr (!kMesh) pair = next_uncertain(scene(), sk, d, o, t, dt, ft, q2);
      } else if (phase == 2 && !do_eval) {
        while (true) {
          if (!(t < ft)) {
            finish(false, q);
            break;
          }
          if (sk.kind == 1 && sp && !(t >= sk.ev0)) {
            // Sphere mean, positive origin, t before the widened outer ball: every
            // grid point up to ev0 lies outside that ball, so it is certain
            // positive (the origin's sign) and the ray still approaches the
            // centre (no early end).  They are stepped with the march's own f32
            // additions, unclassified.  A ray that misses the ball (ev0 = inf)
            // never changes sign: a miss, as the full march finds.
            if (!(sk.ev0 < INFINITY)) {
              finish(false, q);
              break;
            }
            const float te = fminf(sk.ev0, ft);
            int k = 0;
            bool stuck = false;
            while (t < te) {
              ++k;
              if (!(t + dt > t)) {
                stuck = true;
                break;
              }
              t += dt;
            }
            nsteps += k;
            gi += k;
            if (stuck) {
              finish(false, q);
              break;
            }
            continue;
          }
          const F3 p = madd(d, t, o);
          if (sk.kind == 1 && uncertain_t(sk, t)) {
            q = p;  // uncertain for any classification: evaluate
            do_eval = true;
            if constexpr (!kMesh) pair = next_uncertain(scene(), sk, d, o, t, dt, ft, q2);
            break;
          }
          bool fin;
          int ins = -1;
          if (kMesh && sk.kind == 2) ins = inside_here();
          const int cs = classify_point<kMesh>(scene(), sk, p, d, &fin, ins);
          if (cs == 0) {
            q = p;
            qins = ins;
            do_eval = true;
            if constexpr (!kMesh) pair = next_uncertain(scene(), sk, d, o, t, dt, ft, q2);
            break;
          }
          ++nsteps;
          if ((cs > 0) != sp) {
            finish(true, p);
            break;
          }
          if (fin || !(t + dt > t)) {  // no later sign change / the grid cannot advance
            finish(false, p);
            break;
          }
          t += dt;
          ++gi;
        }
      }
    }
    if (__ballot(phase != 3) == 0) break;
    const bool cond_pass = kFuse && phase == 7;
    const bool norm_pass = kFuse && phase == 6;
    if (cond_pass) q = c.C;
    if (norm_pass) q = madd(d, t, o);  // the hit point

    // Window: cells entering the 27-cell neighbourhood of the evaluation point.
    int newcells = -1, nnew = 0;
    int cx = 0, cy = 0, cz = 0;
    if (do_eval || cond_pass || norm_pass) {
      cell_coords(q, G, cc.inv_kr, cc.kr_pow2, cx, cy, cz);
      if (pair) {
        int px2, py2, pz2;
        cell_coords(q2, G, cc.inv_kr, cc.kr_pow2, px2, py2, pz2);
        pair = (px2 == cx) & (py2 == cy) & (pz2 == cz);
      }
      if (!wvalid) {
        newcells = 27;
        nnew = 27;
      } else {
        const int ex = cx - wx, ey = cy - wy, ez = cz - wz;
        if (ex | ey | ez) {
          const bool adj = abs(ex) <= 1 && abs(ey) <= 1 && abs(ez) <= 1;
          newcells = adj ? (ex + 1) * 9 + (ey + 1) * 3 + (ez + 1) : 27;
          nnew = adj ? new_cell_count(ex, ey, ez) : 27;
        }
      }
    }
    // Regeneration of the entered cells of every group of the wave, as one pool
    // over the 64 lanes, written to the same window slots a group's own lanes
    // would write.
    const int items = nnew * ppc;
    if (gl == 0 && newcells >= 0) {
      regen_cell[grp] = make_int4(cx, cy, cz, (int)seed);
      regen_aux[grp] = make_int2(newcells | mod3(cx) << 5 | mod3(cy) << 7 | mod3(cz) << 9 | ppc << 11,
                                 __float_as_int(G.kr));
    }
    __syncthreads();  // the previous evaluation's reads of the window are complete
    int start[NG + 1];  // first pooled item (batched) or entered cell of each group (wave-uniform)
    start[0] = 0;
    if constexpr (kMulti) {
      // item j of group g: impulse i of entered cell k, with the group's own ppc
#pragma unroll
      for (int g = 0; g < NG; ++g) start[g + 1] = start[g] + __builtin_amdgcn_readlane(items, g * GS);
      for (int j = lane; j < start[NG]; j += 64) {
        int g = 0;
#pragma unroll
        for (int h = 1; h < NG; ++h) g += j >= start[h] ? 1 : 0;
        const int4 rc = regen_cell[g];
        const int2 ra = regen_aux[g];
        const int gppc = ra.x >> 11;
        const int it = j - start[g];
        const int k = it / gppc;
        const int i = it - k * gppc;
        const int no = kNewCells.off[ra.x & 31][k];
        const int ox = (no & 3) - 1, oy = ((no >> 2) & 3) - 1, oz = (no >> 4) - 1;  // in {-1,0,1}
        int rx = ((ra.x >> 5) & 3) + ox, ry = ((ra.x >> 7) & 3) + oy, rz = ((ra.x >> 9) & 3) + oz;
        rx = rx < 0 ? rx + 3 : (rx > 2 ? rx - 3 : rx);
        ry = ry < 0 ? ry + 3 : (ry > 2 ? ry - 3 : ry);
        rz = rz < 0 ? rz + 3 : (rz > 2 ? rz - 3 : rz);
        win[g][(rx * 9 + ry * 3 + rz) * gppc + i] = make_impulse(
            __int_as_float(ra.y), rc.x + ox, rc.y + oy, rc.z + oz, (uint32_t)rc.w, kJump.a[i],
            kJump.c[i]);
      }
    } else {
      // Two steps, 64 entered cells at a time: one lane per cell decodes it once
      // (coordinates, first PCG state make_seed + 1, base position n kr, window
      // slot); then one lane per impulse i takes the cell's values by a lane
      // permute and its four draws from states 4i .. 4i + 3 of the cell's stream
      // (make_impulse's draws, computed without the chain of steps).
#pragma unroll
      for (int g = 0; g < NG; ++g) start[g + 1] = start[g] + __builtin_amdgcn_readlane(nnew, g * GS);
      const int ncell = start[NG];
      float4* const wflat = &win[0][0];
      for (int c0 = 0; c0 < ncell; c0 += 64) {
        const int cl = c0 + lane;
        uint32_t s0 = 0u;
        float bx = 0.0f, by = 0.0f, bz = 0.0f;
        int slot = 0;
        if (cl < ncell) {
          int g = 0;
#pragma unroll
          for (int h = 1; h < NG; ++h) g += cl >= start[h] ? 1 : 0;
          const int4 rc = regen_cell[g];
          const int2 ra = regen_aux[g];
          const int k = cl - start[g];
          const int no = kLdsTables ? new_off[(ra.x & 31) * 27 + k] : kNewCells.off[ra.x & 31][k];
          const int ox = (no & 3) - 1, oy = ((no >> 2) & 3) - 1, oz = (no >> 4) - 1;  // in {-1,0,1}
          int rx = ((ra.x >> 5) & 3) + ox, ry = ((ra.x >> 7) & 3) + oy, rz = ((ra.x >> 9) & 3) + oz;
          rx = rx < 0 ? rx + 3 : (rx > 2 ? rx - 3 : rx);
          ry = ry < 0 ? ry + 3 : (ry > 2 ? ry - 3 : ry);
          rz = rz < 0 ? rz + 3 : (rz > 2 ? rz - 3 : rz);
          const int nx = rc.x + ox, ny = rc.y + oy, nz = rc.z + oz;
          s0 = make_seed4((uint32_t)nx, (uint32_t)ny, (uint32_t)nz, (uint32_t)rc.w) + 1u;
          bx = (float)nx * G.kr;
          by = (float)ny * G.kr;
          bz = (float)nz * G.kr;
          slot = g * (27 * PPC) + (rx * 9 + ry * 3 + rz) * PPC;
        }
        const int nitem = min(64, ncell - c0) * PPC;
        for (int m = 0; m < nitem; m += 64) {  // every lane takes part in the permutes
          const int j = m + lane;
          const int c = div_small<PPC>(j);  // j < 64 PPC: nitem <= 64 PPC
          const int src = min(c, 63) * 4;
          const uint32_t cs0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)s0);
          const float ex = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(bx)));
          const float ey = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(by)));
          const float ez = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(bz)));
          const int sl = __builtin_amdgcn_ds_bpermute(src, slot);
          if (j < nitem) {
            const int i = j - c * PPC;
            uint32_t r[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) {
              const int k = 4 * i + d;
              const ulonglong2 st = kLdsTables ? states[k] : make_ulonglong2(kStates.a[k], kStates.c[k]);
              r[d] = pcg_draw_from(cs0, st.x, st.y);
            }
            wflat[sl + i] = make_float4(fmaf(u01(r[2]), G.kr, ex), fmaf(u01(r[1]), G.kr, ey),
                                        fmaf(u01(r[0]), G.kr, ez), (r[3] >> 31) ? 1.0f : -1.0f);
          }
        }
      }
    }
    if (newcells >= 0) {
      wx = cx;
      wy = cy;
      wz = cz;
      wvalid = true;
    }
    __syncthreads();
    if constexpr (kMesh) {
      // Mesh means with cells of more than two grid steps: a bound of |psi| over
      // every point of the regenerated window's cell.  psi there sums exactly
      // these 27 x ppc impulses (weights +-1), each term at most exp(-dist^2 / L^2)
      // for the impulse's distance to the cell's box (widened for the rounding of
      // the cell test, the exp2 and the sum).
      const bool wb_now = newcells >= 0 && G.kr > 2.0f * dt;
      if (__ballot(wb_now) != 0) {
        float part = 0.0f;
        if (wb_now) {
          float blo[3], bhi[3];
          const int wc[3] = {wx, wy, wz};
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            const int nc = wc[ax];
            const float lo_c = (float)(nc > 0 ? nc : nc - 1) * G.kr;
            const float hi_c = (float)(nc >= 0 ? nc + 1 : nc) * G.kr;
            const float wd = 1e-5f * (fabsf(lo_c) + fabsf(hi_c) + G.kr);
            blo[ax] = lo_c - wd;
            bhi[ax] = hi_c + wd;
          }
          for (int sl = gl; sl < 27 * ppc; sl += GS) {
            const float4 im = W[sl];
            const float ex = fmaxf(fmaxf(blo[0] - im.x, im.x - bhi[0]), 0.0f);
            const float ey = fmaxf(fmaxf(blo[1] - im.y, im.y - bhi[1]), 0.0f);
            const float ez = fmaxf(fmaxf(blo[2] - im.z, im.z - bhi[2]), 0.0f);
            const float d2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
            part += __builtin_amdgcn_exp2f(d2 * 0.9999f * cc.psi_k);
          }
        }
        const float tot = group_sum<GS>(part);
        if (wb_now) wbound = tot * fabsf(cc.psi_scale) * 1.001f + 1e-6f;
      }
    }

    // Window sums at q (and q2): each lane takes the same slots in the same order
    // for either point, so a point's sum does not depend on whether it was paired.
    const bool any_pair = !kMesh && __ballot(pair) != 0;
    if (!pair) q2 = q;
    float acc = 0.0f, acc2 = 0.0f;
    // sphere means: every group sums (an idle group's result is not used), which
    // saves the exec-mask bookkeeping; mesh means keep their BVH work guarded
    if (!kMesh || do_eval) {
      if constexpr (kMulti) {
        switch (ppc) {  // the compile-time slot count, per ray
          case 1: window_sums<1, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
          case 2: window_sums<2, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
          case 3: window_sums<3, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
          case 4: window_sums<4, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
          case 5: window_sums<5, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
          case 6: window_sums<6, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
          case 7: window_sums<7, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
          default: window_sums<8, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2); break;
        }
      } else {
        window_sums<PPC, GS>(W, q, q2, any_pair, gl, cc.psi_k, acc, acc2);
      }
    }
    // lanes gl < GS/2 finish q's sum, the others q2's (the same total as
    // group_sum gives, addition being commutative)
    const bool lo = gl < GS / 2;
    acc = any_pair ? group_sum2<GS>(lo ? acc : acc2, lo ? acc2 : acc) : group_sum<GS>(acc);

    const F3 qm = lo ? q : q2;
    float fm = 0.0f;
    if (!kMesh || do_eval) fm = cached_field<kMesh>(scene(), G, qm, acc * cc.psi_scale, c, cc.psi_k, qins);
    float f = fm, f2 = fm;
    if (any_pair) {
      const float fo = swizzle_xor<GS / 2>(fm);
      f = lo ? fm : fo;
      f2 = lo ? fo : fm;
    }
    if (do_eval) {
      if (gl == 0) evals += pair ? 2 : 1;
      if (phase == 1) {
        fprev = f;
        sp = f > 0.0f;
        t = nt;
        gi = 0;
        gi0 = -1;
        phase = 2;
      } else {
        ++nsteps;
        if ((f > 0.0f) != sp && fabsf(fprev - f) > 1e-6f) {
          finish(true, q);
        } else if (!(t + dt > t)) {
          finish(false, q);
        } else {
          float tl = t, fl = f;  // the last evaluated grid point
          t += dt;
          ++gi;
          if (pair) {  // the next grid point, as the next pass would take it
            ++nsteps;
            if ((f2 > 0.0f) != sp && fabsf(fprev - f2) > 1e-6f) {
              finish(true, q2);
            } else if (!(t + dt > t)) {
              finish(false, q2);
            } else {
              tl = t;
              fl = f2;
              t += dt;
            }
          }
          // Sphere means: the grid points after the last evaluated one that stay in
          // its cell (the same 27-cell window, so f is Lipschitz there) and lie
          // within |f| / lip of it keep f's sign, the origin's: no sign change,
          // stepped with the march's own f32 additions.
          if (phase == 2 && sk.kind == 1 && (fl > 0.0f) == sp) {
            const float margin =
                0.01f + 1e-3f * fabsf(fl) +
                lip * 1e-5f * (fabsf(o.x) + fabsf(o.y) + fabsf(o.z) + tl);
            const float reach = (fabsf(fl) - margin) * inv_lip;
            // points before the cell's exit (less a rounding margin) and within reach
            const float tfast = fminf(fminf(ft, tl + reach), cell_exit_t(o, d, wx, wy, wz, G.kr));
            while (t < tfast) {
              ++nsteps;
              const float tn = t + dt;
              if (!(tn > t)) {
                finish(false, o);
                break;
              }
              t = tn;
            }
            // the rest, each checked against the cell as it is computed
            while (phase == 2 && t < ft && t - tl < reach) {
              const F3 p = madd(d, t, o);
              int px2, py2, pz2;
              cell_coords(p, G, cc.inv_kr, cc.kr_pow2, px2, py2, pz2);
              if ((px2 != wx) | (py2 != wy) | (pz2 != wz)) break;
              ++nsteps;
              ++gi;
              if (!(t + dt > t)) {
                finish(false, p);
                break;
              }
              t += dt;
            }
          }
          if constexpr (kMesh) {
            // Mesh means: the grid points after the last evaluated one up to the
            // window cell's exit (less a rounding margin) are certain positive,
            // the origin's sign, when that segment lies outside the +x parity slab
            // (inside() is exactly false there, so the mean is the distance), the
            // ray moves away from the mesh's box (the distance to a box is convex
            // along a line, so it only grows) and the distance at the segment's
            // start exceeds the cell's psi bound plus the conditioning bound with
            // cached_field's margin: stepped with the march's own f32 additions.
            const DevMesh& M = scene().mean_mesh;
            if (phase == 2 && sk.kind == 2 && sp && wbound < INFINITY && M.xf.s > 0.0f) {
              const float te = fminf(ft, cell_exit_t(o, d, wx, wy, wz, G.kr));
              if (t < te) {
                const F3 pa = xf_to_local(M.xf, madd(d, t, o));
                const F3 pb = xf_to_local(M.xf, madd(d, te, o));
                const float ea = 1e-4f * (1.0f + fmaxf(fabsf(pa.x) + fabsf(pa.y) + fabsf(pa.z),
                                                       fabsf(pb.x) + fabsf(pb.y) + fabsf(pb.z)));
                const bool out_slab =
                    (fminf(pa.x, pb.x) > M.lbmax[0] + ea) || (fminf(pa.y, pb.y) > M.lbmax[1] + ea) ||
                    (fmaxf(pa.y, pb.y) < M.lbmin[1] - ea) || (fminf(pa.z, pb.z) > M.lbmax[2] + ea) ||
                    (fmaxf(pa.z, pb.z) < M.lbmin[2] - ea);
                const float pv[3] = {pa.x, pa.y, pa.z}, dv[3] = {d.x, d.y, d.z};
                float e2 = 0.0f, grow = 0.0f;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                  const float over = pv[ax] - M.lbmax[ax], under = M.lbmin[ax] - pv[ax];
                  if (over > 0.0f) {
                    e2 = fmaf(over, over, e2);
                    grow = fmaf(over, dv[ax], grow);
                  } else if (under > 0.0f) {
                    e2 = fmaf(under, under, e2);
                    grow = fmaf(-under, dv[ax], grow);
                  }
                }
                const float low = sqrtf(e2) * M.xf.s * (1.0f - 1e-5f) - ea * M.xf.s;
                const float eb = wbound + condb;
                if (out_slab && grow >= 0.0f && low > eb + 1e-3f + 1e-5f * (eb + low)) {
                  int k = 0;
                  while (t < te) {
                    ++k;
                    const float tn = t + dt;
                    if (!(tn > t)) {
                      finish(false, o);
                      break;
                    }
                    t = tn;
                  }
                  nsteps += k;
                  gi += k;
                }
              }
            }
          }
        }
      }
    }
    if constexpr (kFuse) {
      // Medium hits shaded here, as the shading kernel's handleVolume does for a
      // scene without emitters (trace.cpp:77-116; Renewal+, gpmedium.cpp:133-162):
      // the hit's normal (psi at the hit point, this realization) and the
      // conditioning (psi at C = the hit point, the next realization) take psi
      // and its gradient from the group's window, summed in psi_eval's order, so
      // every value is the shading kernel's.  Each lane first replaces its own
      // slots (the ones it summed) by their terms; after the barrier every lane
      // of the group adds the 27 x PPC terms in that order.
      bool hit_now = norm_pass;
      if (fuse && phase == 5) {
        int hx, hy, hz;
        cell_coords(madd(d, t, o), G, cc.inv_kr, cc.kr_pow2, hx, hy, hz);
        if (wvalid && hx == wx && hy == wy && hz == wz) hit_now = true;
        else phase = 6;  // the window moves to the hit point's cell first
      }
      if (__ballot(hit_now || cond_pass) != 0) {
        const F3 at = hit_now ? madd(d, t, o) : c.C;  // the hit point, or C
        // the path state the hit's shading reads, loaded before the sums below
        uint32_t fl = 0u;
        Pcg rng{0ull};
        F3 T = mk(0.0f, 0.0f, 0.0f);
        if (hit_now) {
          fl = MARCH_ARG(flags)[rid];
          rng.s = MARCH_ARG(rng)[rid];
          T = mk(MARCH_ARG(Tx)[rid], MARCH_ARG(Ty)[rid], MARCH_ARG(Tz)[rid]);
        }
        if (hit_now || cond_pass) {
          for (int sl = gl; sl < 27 * PPC; sl += GS) {
            const float4 im = W[sl];
            W[sl] = psi_terms(G, at, im.x, im.y, im.z, im.w);
          }
        }
        __syncthreads();
        if (hit_now || cond_pass) {
          float sv = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
          const int r0x = mod3(wx - 1), r0y = mod3(wy - 1), r0z = mod3(wz - 1);
#pragma unroll 1
          for (int ax = 0; ax < 3; ++ax) {  // neighbour dx = ax - 1, window residue rx
            const int rx = r0x + ax < 3 ? r0x + ax : r0x + ax - 3;
#pragma unroll 1
            for (int ay = 0; ay < 3; ++ay) {
              const int ry = r0y + ay < 3 ? r0y + ay : r0y + ay - 3;
#pragma unroll 1
              for (int az = 0; az < 3; ++az) {
                const int rz = r0z + az < 3 ? r0z + az : r0z + az - 3;
                const float4* cw = W + (rx * 9 + ry * 3 + rz) * PPC;
#pragma unroll
                for (int i = 0; i < PPC; ++i) {
                  const float4 tm = cw[i];
                  sv = sv + tm.x;
                  sx = sx + tm.y;
                  sy = sy + tm.z;
                  sz = sz + tm.w;
                }
              }
            }
          }
          float pv;
          F3 pg, mg;
          psi_finish(G, sv, sx, sy, sz, pv, pg);
          const float mu = mean_eval<true, false>(scene(), at, &mg);
          wvalid = false;  // the window holds terms now
          if (hit_now) {
            F3 gr;
            field_combine<true>(G, at, c, mu, mg, pv, pg, &gr);
            const F3 nrm = normalized(gr);
            F3 wi, wt;
            bsdf_sample(MUPSI_BSDF_LAMBERTIAN, mk(0.8f, 0.8f, 0.8f), neg(d), nrm, rng, wi, wt);
            T = cmul(T, wt);
            const float surv = max3(T);
            const int bounce = (int)(fl & kBounceMask) + 1;
            if (gl == 0) refsteps += (uint32_t)nsteps;
            if (rng.next1D() > surv || bounce >= scene().max_bounce) {
              // the path ends here (roulette, or its last bounce); its radiance
              // stays as it is and the shading kernel passes it over
              if (gl == 0) const_cast<uint8_t*>(MARCH_ARG(hit))[rid] = 2;
              phase = 0;
            } else {
              T = divs(T, surv);
              const bool single = scene().gp_mode != MUPSI_GP_RENEWAL_PLUS;
              const PathSrc src{MARCH_ARG(src_W),   MARCH_ARG(src_npix), MARCH_ARG(src_k_begin),
                                MARCH_ARG(src_k_begin_dev), MARCH_ARG(src_xs), MARCH_ARG(src_ys),
                                MARCH_ARG(src_ks), nullptr, 0, nullptr};
              int x, y, k;
              path_xyk(src, rid, x, y, k);
              const uint32_t ns =
                  single ? scene().single_seed_i
                         : constant_sampler_i(make_seed5((uint32_t)x, (uint32_t)y, (uint32_t)k, 42u,
                                                         (uint32_t)bounce));
              const Ray nr = make_ray(madd(nrm, scene().eps, at), wi);
              if (gl == 0) {  // the path state the shading kernel would leave
                MARCH_ARG(rng)[rid] = rng.s;
                MARCH_ARG(Tx)[rid] = T.x;
                MARCH_ARG(Ty)[rid] = T.y;
                MARCH_ARG(Tz)[rid] = T.z;
                const_cast<float*>(MARCH_ARG(ox))[rid] = nr.o.x;
                const_cast<float*>(MARCH_ARG(oy))[rid] = nr.o.y;
                const_cast<float*>(MARCH_ARG(oz))[rid] = nr.o.z;
                const_cast<float*>(MARCH_ARG(dx))[rid] = nr.d.x;
                const_cast<float*>(MARCH_ARG(dy))[rid] = nr.d.y;
                const_cast<float*>(MARCH_ARG(dz))[rid] = nr.d.z;
                const_cast<float*>(MARCH_ARG(nt))[rid] = scene().near_eps;
                const_cast<uint32_t*>(MARCH_ARG(flags))[rid] =
                    (uint32_t)bounce | kFlagMedium | (single ? 0u : kFlagCond) | (fl & kRecMask);
                const_cast<uint32_t*>(MARCH_ARG(seed))[rid] = ns;
              }
              o = nr.o;
              d = nr.d;
              nt = scene().near_eps;
              ft = 2000.0f;  // min(FLT_MAX, 2000): render rays have no far bound
              seed = ns;
              c.active = false;
              if (single) {
                next_ray = true;
                phase = 0;
              } else {  // condition at C = the hit point with normal nrm first
                c.C = at;
                c.g = nrm;
                phase = 7;
              }
            }
          } else {
            double u, g3[3];
            renewal_solve_at(G, c.g, mu, mg, pv, pg, u, g3);
            c.active = true;
            c.u = (float)u;
            c.g = mk((float)g3[0], (float)g3[1], (float)g3[2]);
            if (gl == 0) {
              const_cast<float*>(MARCH_ARG(cx))[rid] = c.C.x;
              const_cast<float*>(MARCH_ARG(cy))[rid] = c.C.y;
              const_cast<float*>(MARCH_ARG(cz))[rid] = c.C.z;
              const_cast<float*>(MARCH_ARG(cu))[rid] = c.u;
              const_cast<float*>(MARCH_ARG(cgx))[rid] = c.g.x;
              const_cast<float*>(MARCH_ARG(cgy))[rid] = c.g.y;
              const_cast<float*>(MARCH_ARG(cgz))[rid] = c.g.z;
            }
            next_ray = true;
            phase = 0;
          }
        }
      }
    }
    if (phase == 4 || phase == 5) {  // retire the ray: outputs at the final t, p = o + d t
      if (gl == 0) {
        const bool hit = phase == 5;
        MARCH_ARG(hit)[rid] = hit ? 1 : 0;
        if (hit) {
          const F3 p = madd(d, t, o);
          MARCH_ARG(t)[rid] = t;
          MARCH_ARG(px)[rid] = p.x;
          MARCH_ARG(py)[rid] = p.y;
          MARCH_ARG(pz)[rid] = p.z;
        }
        int32_t* const steps = MARCH_ARG(steps);
        if (steps) steps[rid] = hit ? nsteps : -1;
        refsteps += hit ? (uint32_t)nsteps
                        : nsteps + (uint32_t)fmaxf(0.0f, ceilf((ft - t) / dt));
      }
      phase = 0;
    }
  }
  const unsigned long long evals_w = wave_sum((unsigned long long)evals);
  const unsigned long long refsteps_w = wave_sum((unsigned long long)refsteps);
  const unsigned long long rays_w = wave_sum((unsigned long long)rays);
  unsigned long long* const stats = MARCH_ARG(stats);
  if (lane == 0 && stats) {
    atomicAdd(&stats[0], evals_w);
    atomicAdd(&stats[1], refsteps_w);
    atomicAdd(&stats[2], rays_w);
  }
}

// The window marcher's kernels.  MarchArgs stays the first parameter: MARCH_ARG
// reads it at offset 0 of the kernel arguments.
// Mean meshes and batched launches.
template <int PPC, bool kMesh, bool kMulti, int GS = kGroup>
__global__ __launch_bounds__(64) void k_march_window(MarchArgs a, DevScene S0, CachedConsts cc0,
                                                     const CachedConsts* ccs) {
  march_window<PPC, kMesh, kMulti, GS>(a, S0, cc0, ccs);
}
// Sphere means of a per-scene launch, which also shade their medium hits
// (MarchArgs::fuse): 5 waves per SIMD, as the marching loop alone needs; the
// registers of the rarely run shading code are spilled there instead.
template <int PPC, int GS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_march_sphere(
    MarchArgs a, DevScene S0, CachedConsts cc0, const CachedConsts* ccs) {
  march_window<PPC, false, false, GS>(a, S0, cc0, ccs);
}

// ====================================================================== kernels: wavefront

struct InitArgs {
  PathSoA P;
  PathSrc src;
  int n;
  int* qA;
  const DevScene* scenes;  // batched launch
};
template <bool kMulti>
__global__ __launch_bounds__(256) void k_init(DevScene S0, InitArgs a) {
  const int pid = blockIdx.x * blockDim.x + threadIdx.x;
  const PathSoA& P = a.P;
  int x, y, k;
  if constexpr (kMulti) {
    if (pid >= *a.src.total) return;
    int lo = 0, hi = a.src.nplan - 1;  // the last plan entry with off <= pid
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.src.plan[mid].off <= pid) lo = mid;
      else hi = mid - 1;
    }
    P.sid[pid] = (uint16_t)lo;
    plan_xyk(a.src.plan[lo], pid, x, y, k);
  } else {
    if (pid >= a.n) return;
    path_xyk(a.src, pid, x, y, k);
  }
  const DevScene& S = pick_scene<kMulti>(S0, a.scenes, P.sid, pid);
  P.rng[pid] = pcg_ctor_state(make_seed4((uint32_t)x, (uint32_t)y, (uint32_t)k, 42u));
  const F3 d = normalized(camera_dir(S, x, y));
  P.ox[pid] = S.cam_pos[0];
  P.oy[pid] = S.cam_pos[1];
  P.oz[pid] = S.cam_pos[2];
  P.dx[pid] = d.x;
  P.dy[pid] = d.y;
  P.dz[pid] = d.z;
  P.nt[pid] = 0.0f;
  P.mis[pid] = 1.0f;
  P.Tx[pid] = 1.0f;
  P.Ty[pid] = 1.0f;
  P.Tz[pid] = 1.0f;
  P.Lx[pid] = 0.0f;
  P.Ly[pid] = 0.0f;
  P.Lz[pid] = 0.0f;
  P.flags[pid] = 0u;
  a.qA[pid] = pid;
}

// Guided sampling at GP-medium hits, and the training records of a guided vertex:
// rec[(pid * K + j) * 13 ..] = sampled direction (3), its pdf q, the path
// radiance after the vertex's light sample (3), the throughput after the vertex's
// weight and Russian roulette (3), and the path radiance when the next
// scattering vertex is reached (3; null and mirror surfaces pass light through,
// an emitter's own emission there still counts).  The difference of the two
// radiances over the throughput is the radiance that arrived directly along the
// sampled direction.
constexpr int kRecFloats = 13;
struct GuideShade {
  int on;
  float rr_boost;    // guided roulette: survive with min(1, rr_boost max(T))
  const float* mix;  // kMix floats per path, written by the inference of this bounce
  int record;
  int K;
  float* rec;
};

struct ShadeArgs {
  PathSoA P;
  PathSrc src;
  const int* qA;
  const int* nA;
  int* qA2;
  int* nA2;
  int* qM2;
  int* nM2;
  GuideShade gs;
  int defer_nee;  // medium light samples go to queue S instead of a serial march
  ShadowSoA SH;
  int* qS;
  int* nS;
  const DevScene* scenes;  // batched launch: per-scene constants; medium paths of
  int* qMm2;               // scenes with ppc > 8 go to queue Mm (lane-per-ray marcher)
  int* nMm2;
  int* qMs2;               // and those of mesh-mean scenes with ppc <= 8 to queue Ms
  int* nMs2;
  // the mean and its gradient (mu, grad) at each medium hit of a mesh-mean
  // scene, computed by k_hit_mean after the march; else null
  const float4* hmean;
};

// The guided tracer keeps T as the plain tracer's throughput (albedo weights and
// the reference's Russian roulette on max(T), trace.cpp:68-72,106-111) and the
// product of its MIS factors (cos/pi)/q in `mis`; radiance is added with T * mis.
// Roulette on T * mis would kill a path in proportion to 1/q -- exactly undoing
// the higher probability with which the guide drew it.  Roulette on T alone
// does not depend on the sampled direction, so the estimator stays unbiased.
// Any survival probability that does not depend on the sample keeps it so; the
// guided tracer may raise it to min(1, rr_boost max(T)) to trade time for noise.
__device__ __forceinline__ float rr_survival(const GuideShade& gs, F3 T) {
  const float m = max3(T);
  return gs.on ? fminf(1.0f, gs.rr_boost * m) : m;
}
// One iteration of the bounce loop of PathTracer::trace (trace.cpp:138-210) after
// the march: handleVolume + sampleCondition for medium hits, else
// Scene::intersect + handleSurface + medium switch, or the skydrome.
// kTri false: no scene of the launch holds a triangle mesh (as a primitive or as
// its mean), so no BVH code is compiled.
template <bool kMulti, bool kTri>
__device__ __forceinline__ void shade_path(const DevScene& S0, const ShadeArgs& a, int qi) {
  const int pid = a.qA[qi];
  const PathSoA& P = a.P;
  const DevScene& S = pick_scene<kMulti>(S0, a.scenes, P.sid, pid);
  int x, y, k;
  path_xyk(a.src, pid, x, y, k, P.sid);
  uint32_t fl = P.flags[pid];
  int bounce = (int)(fl & kBounceMask);
  bool medium = (fl & kFlagMedium) != 0;
  Cond cur;
  cur.active = (fl & kFlagCond) != 0;
  if (cur.active) {
    cur.C = mk(P.cx[pid], P.cy[pid], P.cz[pid]);
    cur.u = P.cu[pid];
    cur.g = mk(P.cgx[pid], P.cgy[pid], P.cgz[pid]);
  }
  Pcg rng{P.rng[pid]};
  F3 T = mk(P.Tx[pid], P.Ty[pid], P.Tz[pid]);
  F3 L = mk(P.Lx[pid], P.Ly[pid], P.Lz[pid]);
  float mis = a.gs.on ? P.mis[pid] : 1.0f;
  auto Te = [&]() { return a.gs.on ? mul(T, mis) : T; };  // throughput radiance is added with
  Ray ray{mk(P.ox[pid], P.oy[pid], P.oz[pid]), mk(P.dx[pid], P.dy[pid], P.dz[pid]), P.nt[pid],
          3.402823466e+38f};
  const bool single = S.gp_mode != MUPSI_GP_RENEWAL_PLUS;
  bool alive = true, hasHit = false, exited = true;
  bool shadow = false;   // a deferred light sample waits in queue S
  bool rec_now = false;  // this event is a guided vertex to record
  F3 rec_wi = mk(0, 0, 0);
  float rec_q = 0.0f;
  F3 rec_L = mk(0, 0, 0);

  uint32_t nrec = (fl & kRecMask) >> kRecShift;
  bool open = (fl & kFlagOpen) != 0;
  auto close_record = [&](F3 Lc) {
    if (a.gs.record && open && nrec > 0) {
      float* r = a.gs.rec + ((size_t)pid * a.gs.K + (nrec - 1)) * kRecFloats;
      r[10] = Lc.x;
      r[11] = Lc.y;
      r[12] = Lc.z;
    }
    open = false;
  };

  // march result: 0 no hit, 1 hit, 2 the marcher shaded the path's medium hits
  // and the path ended there (roulette or its last bounce)
  const uint8_t mh = (medium && !S.skip_medium) ? P.mhit[pid] : 0;
  if (mh == 2) alive = false;
  if (mh == 1) {
    // handleVolume (trace.cpp:77-116)
    close_record(L);
    const uint32_t iseed = P.mseed[pid];
    const F3 Pp = mk(P.mpx[pid], P.mpy[pid], P.mpz[pid]);
    // the mean at the hit serves the normal and the Renewal+ solve at C = Pp
    MeanAt mp{};
    F3 gr;
    if (S.medium_kind == MUPSI_MEDIUM_GP) {
      if (a.hmean && S.mean_type == MUPSI_MEAN_MESH && S.mean_mesh.n_nodes > 0) {
        // mean_eval<true>(S, Pp), computed by k_hit_mean
        const float4 h = a.hmean[pid];
        mp.mu = h.x;
        mp.grad = mk(h.y, h.z, h.w);
      } else {
        mp.mu = mean_eval<true, kTri>(S, Pp, &mp.grad);
      }
      field_eval<true, kTri>(S, Pp, iseed, cur, &gr, &mp);
    } else {
      gr = alt_grad<kTri>(S, Pp, iseed);
    }
    const F3 nrm = normalized(gr);
    const F3 wo = neg(ray.d);
    const F3 alb = mk(0.8f, 0.8f, 0.8f);
    LightSample ls;
    if (choose_light<kTri>(S, Pp, rng, ls)) {
      Ray sr = make_ray(madd(nrm, S.eps, Pp), ls.d);
      const float bias = ls.dist - S.eps * fabsf(dot(ls.d, nrm));
      sr.farT = bias * 0.99f;
      const F3 f = bsdf_weight(MUPSI_BSDF_LAMBERTIAN, alb, wo, nrm, ls.d);
      const float ac = fabsf(dot(ls.d, nrm));
      const F3 Tt = Te();
      if (a.defer_nee) {
        const ShadowSoA& H = a.SH;
        H.ox[pid] = sr.o.x;
        H.oy[pid] = sr.o.y;
        H.oz[pid] = sr.o.z;
        H.dx[pid] = sr.d.x;
        H.dy[pid] = sr.d.y;
        H.dz[pid] = sr.d.z;
        H.nt[pid] = sr.nearT;
        H.ft[pid] = sr.farT;
        H.seed[pid] = iseed;
        H.active[pid] = cur.active ? 1 : 0;
        if (cur.active) {
          H.cx[pid] = cur.C.x;
          H.cy[pid] = cur.C.y;
          H.cz[pid] = cur.C.z;
          H.cu[pid] = cur.u;
          H.cgx[pid] = cur.g.x;
          H.cgy[pid] = cur.g.y;
          H.cgz[pid] = cur.g.z;
        }
        H.pend_x[pid] = Tt.x * (f.x * ls.w.x);
        H.pend_y[pid] = Tt.y * (f.y * ls.w.y);
        H.pend_z[pid] = Tt.z * (f.z * ls.w.z);
        H.ac[pid] = ac;
        shadow = true;
      } else {
        const float trans = march_serial<kTri>(S, sr, iseed, cur) ? 0.0f : 1.0f;
        L = mk(fmaf(trans * (Tt.x * (f.x * ls.w.x)), ac, L.x),
               fmaf(trans * (Tt.y * (f.y * ls.w.y)), ac, L.y),
               fmaf(trans * (Tt.z * (f.z * ls.w.z)), ac, L.z));
      }
    }
    F3 wi, wt;
    float q = 0.0f;
    if (a.gs.on) {
      float m[guide::kMix];
      load_mix(a.gs.mix + (size_t)pid * guide::kMix, m);
      const float u0 = rng.next1D();
      float u1, u2, u3 = 0.0f;
      u1 = rng.next1D();
      u2 = rng.next1D();
      if (u0 < m[guide::kAlphaIdx]) u3 = rng.next1D();
      wi = guided_dir(m, nrm, u0, u1, u2, u3);
      float wfac;
      guided_eval(m, nrm, wi, q, wfac);
      wt = alb;
      mis *= wfac;
    } else {
      bsdf_sample(MUPSI_BSDF_LAMBERTIAN, alb, wo, nrm, rng, wi, wt);
    }
    T = cmul(T, wt);
    const float surv = rr_survival(a.gs, T);
    if ((a.gs.on && !(mis > 0.0f)) || rng.next1D() > surv) {  // mis 0: below the surface
      alive = false;
    } else {
      T = divs(T, surv);
      if (a.gs.record) {
        rec_now = true;
        rec_wi = wi;
        rec_q = q;
        rec_L = L;
      }
      if (!single && S.medium_kind == MUPSI_MEDIUM_GP) {  // only GPMedium re-conditions
        const uint32_t ns =
            constant_sampler_i(make_seed5((uint32_t)x, (uint32_t)y, (uint32_t)k, 42u, (uint32_t)(bounce + 1)));
        double u, g[3];
        renewal_solve(S, Pp, nrm, ns, u, g, &mp);
        cur.active = true;
        cur.C = Pp;
        cur.u = (float)u;
        cur.g = mk((float)g[0], (float)g[1], (float)g[2]);
      } else {
        cur.active = false;
      }
      ray = make_ray(madd(nrm, S.eps, Pp), wi);
      ray.nearT = S.near_eps;
      ++bounce;
      exited = false;
      hasHit = true;
    }
  }

  if (alive && exited) {
    F3 Pp, Ng;
    float uu = 0.f, vv = 0.f;
    const int hp = scene_intersect<kTri>(S, ray, Pp, Ng, uu, vv);
    if (hp >= 0) {
      // handleSurface (trace.cpp:32-75)
      const DevPrim& pr = S.prims[hp];
      const F3 wo = neg(ray.d), nrm = Ng;
      const F3 L0 = L;
      F3 alb = mk(1, 1, 1);
      if (pr.bsdf == MUPSI_BSDF_LAMBERTIAN)
        alb = pr.albedo.rgb ? tex_at(pr.albedo, uu, vv, nullptr) : mk(0.8f, 0.8f, 0.8f);
      LightSample ls;
      if (choose_light<kTri>(S, Pp, rng, ls)) {
        Ray sr = make_ray(madd(nrm, S.eps, Pp), ls.d);
        const float bias = ls.dist - S.eps * fabsf(dot(ls.d, nrm));
        sr.farT = bias * 0.99f;
        if (!scene_occluded<kTri>(S, sr)) {
          const F3 w = bsdf_weight(pr.bsdf, alb, wo, nrm, ls.d);
          const float ac = fabsf(dot(ls.d, nrm));
          const F3 Tt = Te();
          L = mk(fmaf(Tt.x * (w.x * ls.w.x), ac, L.x), fmaf(Tt.y * (w.y * ls.w.y), ac, L.y),
                 fmaf(Tt.z * (w.z * ls.w.z), ac, L.z));
        }
      }
      F3 wi, wt;
      bsdf_sample(pr.bsdf, alb, wo, nrm, rng, wi, wt);
      F3 Lc = L0;
      if (pr.has_em) {
        const float mm = fmaxf(0.0f, dot(nrm, wi));
        const F3 Tt = Te();
        L = mk(fmaf(Tt.x, pr.em[0] * mm, L.x), fmaf(Tt.y, pr.em[1] * mm, L.y),
               fmaf(Tt.z, pr.em[2] * mm, L.z));
        Lc = mk(fmaf(Tt.x, pr.em[0] * mm, L0.x), fmaf(Tt.y, pr.em[1] * mm, L0.y),
                fmaf(Tt.z, pr.em[2] * mm, L0.z));
      }
      if (pr.bsdf == MUPSI_BSDF_LAMBERTIAN) close_record(Lc);
      T = cmul(T, wt);
      const float surv = rr_survival(a.gs, T);
      if (rng.next1D() > surv) {
        alive = false;
      } else {
        T = divs(T, surv);
        const bool back = dot(wi, Ng) < 0.0f;
        medium = back ? (pr.int_medium != 0 && S.has_gp != 0) : false;
        ray = make_ray(madd(mul(Ng, back ? -1.0f : 1.0f), S.eps, Pp), wi);
        ray.nearT = S.near_eps;
        ++bounce;
        hasHit = true;
      }
    } else if (S.sky.rgb) {
      const F3 e = sky_lookup(S, ray.d, nullptr);
      const F3 Tt = Te();
      L = mk(fmaf(Tt.x, e.x, L.x), fmaf(Tt.y, e.y, L.y), fmaf(Tt.z, e.z, L.z));
    }
  }

  const bool cont = alive && hasHit && bounce < S.max_bounce;
  bool rec_wait = false;
  P.Lx[pid] = L.x;
  P.Ly[pid] = L.y;
  P.Lz[pid] = L.z;
  if (cont && rec_now && (int)nrec < a.gs.K) {
    float* r = a.gs.rec + ((size_t)pid * a.gs.K + nrec) * kRecFloats;
    r[0] = rec_wi.x;
    r[1] = rec_wi.y;
    r[2] = rec_wi.z;
    r[3] = rec_q;
    r[4] = rec_L.x;
    r[5] = rec_L.y;
    r[6] = rec_L.z;
    const F3 Tt = Te();  // the throughput radiance after this vertex is added with
    r[7] = Tt.x;
    r[8] = Tt.y;
    r[9] = Tt.z;
    ++nrec;
    open = true;
    rec_wait = shadow;  // r[4..6] still lacks this vertex's light sample
  }
  if (cont) {
    P.rng[pid] = rng.s;
    P.Tx[pid] = T.x;
    P.Ty[pid] = T.y;
    P.Tz[pid] = T.z;
    if (a.gs.on) P.mis[pid] = mis;
    P.ox[pid] = ray.o.x;
    P.oy[pid] = ray.o.y;
    P.oz[pid] = ray.o.z;
    P.dx[pid] = ray.d.x;
    P.dy[pid] = ray.d.y;
    P.dz[pid] = ray.d.z;
    P.nt[pid] = ray.nearT;
    P.flags[pid] = (uint32_t)bounce | (medium ? kFlagMedium : 0u) | (cur.active ? kFlagCond : 0u) |
                   (nrec << kRecShift) | (open ? kFlagOpen : 0u) | (rec_wait ? kFlagNeeRec : 0u);
    if (cur.active) {
      P.cx[pid] = cur.C.x;
      P.cy[pid] = cur.C.y;
      P.cz[pid] = cur.C.z;
      P.cu[pid] = cur.u;
      P.cgx[pid] = cur.g.x;
      P.cgy[pid] = cur.g.y;
      P.cgz[pid] = cur.g.z;
    }
    if (medium && !S.skip_medium) {
      P.mseed[pid] = single ? S.single_seed_i
                            : constant_sampler_i(make_seed5((uint32_t)x, (uint32_t)y, (uint32_t)k,
                                                            42u, (uint32_t)bounce));
    }
  }
  wave_push(cont, pid, a.qA2, a.nA2);
  if constexpr (kMulti) {
    const bool cached = S.medium_kind == MUPSI_MEDIUM_GP && S.gp.ppc <= kMaxPpcCached;
    const bool mesh = S.mean_type == MUPSI_MEAN_MESH;
    const bool go = cont && medium && !S.skip_medium;
    wave_push(go && cached && !mesh, pid, a.qM2, a.nM2);
    wave_push(go && cached && mesh, pid, a.qMs2, a.nMs2);
    wave_push(go && !cached, pid, a.qMm2, a.nMm2);
  } else {
    wave_push(cont && medium && !S.skip_medium, pid, a.qM2, a.nM2);
  }
  if (a.defer_nee) wave_push(shadow, pid, a.qS, a.nS);
}

template <bool kMulti, bool kTri>
__global__ __launch_bounds__(256) void k_shade(DevScene S0, ShadeArgs a) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi < *a.nA) shade_path<kMulti, kTri>(S0, a, qi);
}

// The deferred light samples after their march (k_march / k_march_window over
// queue S): add the contribution where the shadow ray found no sign change, in
// the reference's operation order, before any later event of the path adds to L.
struct NeeArgs {
  PathSoA P;
  ShadowSoA SH;
  const int* qS;
  const int* nS;
  GuideShade gs;
};
__global__ __launch_bounds__(256) void k_nee_finish(NeeArgs a) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= *a.nS) return;
  const int pid = a.qS[qi];
  const ShadowSoA& H = a.SH;
  const PathSoA& P = a.P;
  const float trans = H.hit[pid] ? 0.0f : 1.0f;
  const float ac = H.ac[pid];
  const float px = trans * H.pend_x[pid], py = trans * H.pend_y[pid], pz = trans * H.pend_z[pid];
  P.Lx[pid] = fmaf(px, ac, P.Lx[pid]);
  P.Ly[pid] = fmaf(py, ac, P.Ly[pid]);
  P.Lz[pid] = fmaf(pz, ac, P.Lz[pid]);
  const uint32_t fl = P.flags[pid];
  if (a.gs.record && (fl & kFlagNeeRec)) {
    const uint32_t nrec = (fl & kRecMask) >> kRecShift;
    float* r = a.gs.rec + ((size_t)pid * a.gs.K + (nrec - 1)) * kRecFloats;
    r[4] = fmaf(px, ac, r[4]);
    r[5] = fmaf(py, ac, r[5]);
    r[6] = fmaf(pz, ac, r[6]);
    P.flags[pid] = fl & ~kFlagNeeRec;
  }
}

// Guided vertices of this bounce: every path of queue M whose march hit the
// medium gets its hit normal (the same evaluation k_shade makes), its network
// input features (row gi of feat, gi = its slot in queue G) and, when recording,
// a copy of the features in its training slot (pid * K + j).
struct GuidePrepArgs {
  PathSoA P;
  const int* qM;
  const int* nM;
  int* qG;
  int* nG;
  float* feat;
  int record;
  int K;
  float* rec_feat;
  guide::EncodeBox box;
};
__global__ __launch_bounds__(256) void k_guide_prep(DevScene S, GuidePrepArgs a) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = *a.nM;
  const PathSoA& P = a.P;
  int pid = 0;
  bool act = false;
  if (qi < n) {
    pid = a.qM[qi];
    act = P.mhit[pid] != 0;
  }
  const int gi = wave_slot(act, a.nG);
  if (!act) return;
  const uint32_t fl = P.flags[pid];
  Cond cur;
  cur.active = (fl & kFlagCond) != 0;
  if (cur.active) {
    cur.C = mk(P.cx[pid], P.cy[pid], P.cz[pid]);
    cur.u = P.cu[pid];
    cur.g = mk(P.cgx[pid], P.cgy[pid], P.cgz[pid]);
  }
  const F3 Pp = mk(P.mpx[pid], P.mpy[pid], P.mpz[pid]);
  F3 gr;
  if (S.medium_kind == MUPSI_MEDIUM_GP) field_eval<true>(S, Pp, P.mseed[pid], cur, &gr);
  else gr = alt_grad(S, Pp, P.mseed[pid]);
  const F3 nrm = normalized(gr);
  float f[guide::kIn];
  guide::encode_features(a.box, Pp.x, Pp.y, Pp.z, nrm.x, nrm.y, nrm.z, -P.dx[pid], -P.dy[pid],
                         -P.dz[pid], f);
  float4* dst = reinterpret_cast<float4*>(a.feat + (size_t)gi * guide::kIn);
#pragma unroll
  for (int r = 0; r < guide::kIn / 4; ++r) dst[r] = make_float4(f[4 * r], f[4 * r + 1], f[4 * r + 2], f[4 * r + 3]);
  a.qG[gi] = pid;
  if (a.record) {
    const int j = (int)((fl & kRecMask) >> kRecShift);
    if (j < a.K) {
      float4* rd = reinterpret_cast<float4*>(a.rec_feat + ((size_t)pid * a.K + j) * guide::kIn);
#pragma unroll
      for (int r = 0; r < guide::kIn / 4; ++r) rd[r] = make_float4(f[4 * r], f[4 * r + 1], f[4 * r + 2], f[4 * r + 3]);
    }
  }
}

// Training samples of a finished pass: for every recorded guided vertex, the
// radiance that arrived along its sampled direction is the path radiance added
// after the vertex divided by the throughput after it.  It splits into the part
// collected before the next scattering vertex (direct: sky, emitters, through
// null and mirror surfaces) and the rest (indirect, which carries the noise of
// every later vertex); the training radiance is lum(direct) + min(lum(indirect),
// clamp), and the sample's weight is that times max(cos, 0) / q (the product of
// incident radiance and cosine, over the density the direction was drawn from).
struct CollectArgs {
  PathSoA P;
  int n;
  int K;
  const float* rec;
  const float* rec_feat;
  float* tr_feat;  // ring of tr_cap records; this pass writes from tr_head on
  float* tr_dir;
  float* tr_w;
  int* tr_count;
  int64_t tr_head, tr_cap;
  float indirect_clamp;
};
__global__ __launch_bounds__(256) void k_collect(CollectArgs a) {
  const int pid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool inr = pid < a.n;
  const uint32_t pfl = inr ? a.P.flags[pid] : 0u;
  const int nrec = (int)((pfl & kRecMask) >> kRecShift);
  const bool last_open = (pfl & kFlagOpen) != 0;
  const F3 Le = inr ? mk(a.P.Lx[pid], a.P.Ly[pid], a.P.Lz[pid]) : mk(0, 0, 0);
  for (int j = 0; j < a.K; ++j) {  // uniform trip count: wave_slot needs every lane
    bool ok = inr && j < nrec;
    float w = 0.0f;
    F3 wi = mk(0, 0, 0);
    const size_t slot = (size_t)pid * a.K + j;
    if (ok) {
      const float* r = a.rec + slot * kRecFloats;
      wi = mk(r[0], r[1], r[2]);
      const float q = r[3];
      // radiance collected before the next scattering vertex (the path's end if
      // the record is still open)
      const F3 Lc = (j == nrec - 1 && last_open) ? Le : mk(r[10], r[11], r[12]);
      const F3 Li = mk(r[7] > 0.0f ? (Lc.x - r[4]) / r[7] : 0.0f,
                       r[8] > 0.0f ? (Lc.y - r[5]) / r[8] : 0.0f,
                       r[9] > 0.0f ? (Lc.z - r[6]) / r[9] : 0.0f);
      const F3 Lind = mk(r[7] > 0.0f ? (Le.x - Lc.x) / r[7] : 0.0f,
                         r[8] > 0.0f ? (Le.y - Lc.y) / r[8] : 0.0f,
                         r[9] > 0.0f ? (Le.z - Lc.z) / r[9] : 0.0f);
      const float lind = fmaxf((Lind.x + Lind.y + Lind.z) * (1.0f / 3.0f), 0.0f);
      const float lum = (Li.x + Li.y + Li.z) * (1.0f / 3.0f) + fminf(lind, a.indirect_clamp);
      const float* fe = a.rec_feat + slot * guide::kIn;
      const float cosv = dot(wi, mk(fe[24], fe[25], fe[26]));
      w = q > 0.0f ? lum * fmaxf(cosv, 0.0f) / q : 0.0f;
      ok = w > 0.0f && isfinite(w);
    }
    const int tslot = wave_slot(ok, a.tr_count);
    if (ok) {
      const int64_t t = (a.tr_head + tslot) % a.tr_cap;
      const float4* src = reinterpret_cast<const float4*>(a.rec_feat + slot * guide::kIn);
      float4* dst = reinterpret_cast<float4*>(a.tr_feat + (size_t)t * guide::kIn);
#pragma unroll
      for (int r = 0; r < guide::kIn / 4; ++r) dst[r] = src[r];
      a.tr_dir[3 * (size_t)t] = wi.x;
      a.tr_dir[3 * (size_t)t + 1] = wi.y;
      a.tr_dir[3 * (size_t)t + 2] = wi.z;
      a.tr_w[t] = w;
    }
  }
}

// mupsi_guide_sample: the guided sample of k_shade on explicit uniforms.
__global__ __launch_bounds__(256) void k_guide_sample(const float* mix, const float* nrm,
                                                      const float* u, int64_t n, float* dir,
                                                      float* pdf, float* weight) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float m[guide::kMix];
  load_mix(mix + i * guide::kMix, m);
  const F3 nn = ld3(nrm, i);
  const F3 wi = guided_dir(m, nn, u[4 * i], u[4 * i + 1], u[4 * i + 2], u[4 * i + 3]);
  float q, wfac;
  guided_eval(m, nn, wi, q, wfac);
  dir[3 * i] = wi.x;
  dir[3 * i + 1] = wi.y;
  dir[3 * i + 2] = wi.z;
  pdf[i] = q;
  weight[i] = wfac;
}

// Per-pixel FP64 accumulation over the passes of one launch, in pass order.
__global__ __launch_bounds__(256) void k_accum(const float* Lx, const float* Ly, const float* Lz,
                                               int npix, int kcount, double* acc) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= npix) return;
  double sx = acc[3 * pix], sy = acc[3 * pix + 1], sz = acc[3 * pix + 2];
  for (int kk = 0; kk < kcount; ++kk) {
    const int64_t i = (int64_t)kk * npix + pix;
    sx += (double)Lx[i];
    sy += (double)Ly[i];
    sz += (double)Lz[i];
  }
  acc[3 * pix] = sx;
  acc[3 * pix + 1] = sy;
  acc[3 * pix + 2] = sz;
}

// Batched launch: scene s's share of the launch into its own accumulator.
__global__ __launch_bounds__(256) void k_accum_plan(const float* Lx, const float* Ly,
                                                    const float* Lz, const LaunchPlan* plan, int s,
                                                    double* acc) {
  const LaunchPlan L = plan[s];
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= L.npix || L.kc <= 0) return;
  double sx = acc[3 * pix], sy = acc[3 * pix + 1], sz = acc[3 * pix + 2];
  for (int kk = 0; kk < L.kc; ++kk) {
    const int64_t i = L.off + (int64_t)kk * L.npix + pix;
    sx += (double)Lx[i];
    sy += (double)Ly[i];
    sz += (double)Lz[i];
  }
  acc[3 * pix] = sx;
  acc[3 * pix + 1] = sy;
  acc[3 * pix + 2] = sz;
}

__global__ void k_set_counters(int* c, int n, int a0, int i0) {
  const int i = threadIdx.x;
  if (i < n) c[i] = (i == i0) ? a0 : 0;
}

// ====================================================================== kernels: batch API

__global__ __launch_bounds__(256) void k_psi(DevGP g, const float* px, const float* py,
                                             const float* pz, const uint32_t* seed, int64_t n,
                                             float* out4, float* l1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v, m;
  F3 gr;
  psi_eval<true>(g, mk(px[i], py[i], pz[i]), seed[i], v, gr, l1 ? &m : nullptr);
  out4[4 * i] = v;
  out4[4 * i + 1] = gr.x;
  out4[4 * i + 2] = gr.y;
  out4[4 * i + 3] = gr.z;
  if (l1) l1[i] = m;
}

__global__ __launch_bounds__(256) void k_cell_impulses(DevGP g, const int32_t* cell3,
                                                       const uint32_t* seed, int64_t n,
                                                       uint32_t* cell_seed, float* pos,
                                                       int8_t* w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int cx = cell3[3 * i], cy = cell3[3 * i + 1], cz = cell3[3 * i + 2];
  const uint32_t cs = make_seed4((uint32_t)cx, (uint32_t)cy, (uint32_t)cz, seed[i]);
  cell_seed[i] = cs;
  uint64_t s = (uint32_t)(cs + 1u);
  const float bx = (float)cx * g.kr, by = (float)cy * g.kr, bz = (float)cz * g.kr;
  for (int j = 0; j < g.ppc; ++j) {
    const uint32_t r0 = pcg_next(s), r1 = pcg_next(s), r2 = pcg_next(s), r3 = pcg_next(s);
    const int64_t o = i * g.ppc + j;
    pos[3 * o] = fmaf(u01(r2), g.kr, bx);
    pos[3 * o + 1] = fmaf(u01(r1), g.kr, by);
    pos[3 * o + 2] = fmaf(u01(r0), g.kr, bz);
    w[o] = (r3 >> 31) ? 1 : -1;
  }
}

struct CondSoA {
  const uint8_t* active;
  const float *cx, *cy, *cz, *u, *gx, *gy, *gz;
};
__device__ __forceinline__ Cond load_cond(const CondSoA& c, int64_t i) {
  Cond r;
  r.active = c.active && c.active[i];
  if (r.active) {
    r.C = mk(c.cx[i], c.cy[i], c.cz[i]);
    r.u = c.u[i];
    r.g = mk(c.gx[i], c.gy[i], c.gz[i]);
  }
  return r;
}

__global__ __launch_bounds__(256) void k_field(DevScene S, const float* px, const float* py,
                                               const float* pz, const uint32_t* seed, CondSoA c,
                                               int64_t n, float* out4) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  F3 g;
  const float v = medium_field<true>(S, mk(px[i], py[i], pz[i]), seed[i], load_cond(c, i), &g);
  out4[4 * i] = v;
  out4[4 * i + 1] = g.x;
  out4[4 * i + 2] = g.y;
  out4[4 * i + 3] = g.z;
}

// Hit normal of sampleDistance: normalize(grad f(p)) (gpmedium.cpp:81-82,
// sdfmedium.cpp:52-53, externalgpmedium.cpp:101-102).
__global__ __launch_bounds__(256) void k_hit_normals(DevScene S, const uint8_t* hit,
                                                     const float* p3x, const float* p3y,
                                                     const float* p3z, const uint32_t* seed,
                                                     CondSoA c, int64_t n, float* nrm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  F3 g = mk(0, 0, 0);
  if (hit[i]) {
    medium_field<true>(S, mk(p3x[i], p3y[i], p3z[i]), seed[i], load_cond(c, i), &g);
    g = normalized(g);
  }
  nrm[3 * i] = g.x;
  nrm[3 * i + 1] = g.y;
  nrm[3 * i + 2] = g.z;
}

__global__ void k_interleave3(const float* x, const float* y, const float* z, int64_t n,
                              float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[3 * i] = x[i];
  out[3 * i + 1] = y[i];
  out[3 * i + 2] = z[i];
}

__global__ __launch_bounds__(256) void k_condition(DevScene S, const float* cx, const float* cy,
                                                   const float* cz, const float* nx,
                                                   const float* ny, const float* nz,
                                                   const uint32_t* seed_next, int64_t n,
                                                   double* u, double* g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double uu, gg[3];
  renewal_solve(S, mk(cx[i], cy[i], cz[i]), mk(nx[i], ny[i], nz[i]), seed_next[i], uu, gg);
  u[i] = uu;
  g[3 * i] = gg[0];
  g[3 * i + 1] = gg[1];
  g[3 * i + 2] = gg[2];
}

// Ray(o, d) construction for the batch march: copy o, normalize d.
__global__ void k_normalize_rays(const float* ox, const float* oy, const float* oz,
                                 const float* dx, const float* dy, const float* dz, int64_t n,
                                 float* a, float* b, float* c, float* e, float* f, float* g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const F3 d = normalized(mk(dx[i], dy[i], dz[i]));
  a[i] = ox[i];
  b[i] = oy[i];
  c[i] = oz[i];
  e[i] = d.x;
  f[i] = d.y;
  g[i] = d.z;
}

struct RaySoA {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *nt, *ft;
};
__device__ __forceinline__ Ray load_ray(const RaySoA& r, int64_t i) {
  Ray ray = make_ray(mk(r.ox[i], r.oy[i], r.oz[i]), mk(r.dx[i], r.dy[i], r.dz[i]));
  ray.nearT = r.nt[i];
  ray.farT = r.ft[i];
  return ray;
}

// The closest-point, intersection and distance-grid kernels copy the top
// levels of the tree (the first kTopRecs records, 32 KiB: every query's
// traversal starts there) into LDS once per block and loop over their queries
// (a grid of a few blocks per CU, so that a block's copy serves many queries);
// the other records are read from global memory.
constexpr uint32_t kTopRecs = 511;
__device__ __forceinline__ DevMesh stage_top(const DevMesh& m, BvhRec* lds) {
  DevMesh t = m;
  const uint32_t nt = m.n_recs < kTopRecs ? m.n_recs : kTopRecs;
  const float4* src = reinterpret_cast<const float4*>(m.recs);
  float4* dst = reinterpret_cast<float4*>(lds);
  for (uint32_t i = threadIdx.x; i < 4 * nt; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  t.top = lds;
  t.ntop = nt;
  return t;
}
#define MESH_QUERY_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

__global__ __launch_bounds__(256) void k_mesh_intersect(DevMesh m0, RaySoA rs, int64_t n,
                                                        int32_t* tri, float* t, float* u,
                                                        float* v) {
  __shared__ BvhRec top[kTopRecs];
  const DevMesh m = stage_top(m0, top);
  MESH_QUERY_LOOP(i, n) {
    Ray r = load_ray(rs, i);
    uint32_t ti = 0;
    float uu = 0.f, vv = 0.f;
    const bool h = mesh_intersect<true>(m, r, ti, uu, vv);
    tri[i] = h ? (int32_t)ti : -1;
    t[i] = h ? r.farT : 0.0f;
    u[i] = uu;
    v[i] = vv;
  }
}
__global__ __launch_bounds__(256) void k_mesh_occluded(DevMesh m, RaySoA rs, int64_t n,
                                                       uint8_t* occ) {
  MESH_QUERY_LOOP(i, n) occ[i] = mesh_occluded(m, load_ray(rs, i)) ? 1 : 0;
}
__global__ __launch_bounds__(256) void k_mesh_closest(DevMesh m0, const float* px, const float* py,
                                                      const float* pz, int64_t n, float* dist,
                                                      int32_t* tri, float* cl) {
  __shared__ BvhRec top[kTopRecs];
  const DevMesh m = stage_top(m0, top);
  MESH_QUERY_LOOP(i, n) {
    uint32_t ti = 0;
    F3 o = mk(0, 0, 0);
    // a mean mesh carries a distance grid: its bound only prunes (same results)
    dist[i] = m.dlb ? mesh_dist_grid<true>(m, mk(px[i], py[i], pz[i]), ti, o)
                    : mesh_dist<true>(m, mk(px[i], py[i], pz[i]), ti, o);
    tri[i] = (int32_t)ti;
    cl[3 * i] = o.x;
    cl[3 * i + 1] = o.y;
    cl[3 * i + 2] = o.z;
  }
}
// inside() reads no LDS copy of the tree's top: a mean mesh's query reads its
// leaf bin (about 7 entries on the 327,680-triangle mesh), and a primitive mesh's
// stackless walk tests ~50 boxes, too few to pay for the copy and the resident
// blocks it costs (measured for the walk: 5.2 ms with the copy, 3.9 ms without,
// 4M queries on the 327,680-triangle mesh).
__global__ __launch_bounds__(256) void k_mesh_inside(DevMesh m, const float* px, const float* py,
                                                     const float* pz, int64_t n, uint8_t* in) {
  MESH_QUERY_LOOP(i, n) in[i] = mesh_inside(m, mk(px[i], py[i], pz[i])) ? 1 : 0;
}
// The hits of a march queue put in order of a key before their closest-point
// searches: in batched launches the scene first, then the distance class (the
// distance grids' upper bound at the hit point relative to the mesh's box
// length mx: a search takes longer the farther its point lies from the
// surface), then the Morton index of the point's cell in a grid of 2^cb cubes
// per axis over the mesh's box padded by mx/4 (points outside are clamped to
// its border cells).  A wave's searches then have similar lengths and start
// near each other, so its lanes read the same nodes.  Counting sort: k_hit_key
// counts each key (one atomic per key present in a wave, which also gives each
// lane its rank among the key's hits), k_hit_scan turns the counts into starts,
// k_hit_place writes each path id at start + rank.  The order within a key
// varies from run to run; every hit's result is its own, so the bits do not.
constexpr int kHitClasses = 4;
constexpr int kHitCellBits = 5;             // single-scene launches: 32^3 cells
constexpr int kHitMaxKeys = 1 << 19;        // batched launches: cb shrinks to fit
constexpr int kHitScanThreads = 1024;
constexpr int kHitKeyQuantum = 4 * kHitScanThreads;  // key counts are padded to this
// Cell bits for a launch over n_scenes scenes: the largest cb <= kHitCellBits
// with n_scenes * kHitClasses * 8^cb <= kHitMaxKeys (cb 0 for 65,536 scenes).
inline int hit_cell_bits(int n_scenes) {
  int cb = kHitCellBits;
  while (cb > 0 && (int64_t)n_scenes * kHitClasses * (1ll << (3 * cb)) > kHitMaxKeys) --cb;
  return cb;
}
inline int hit_key_count(int n_scenes, int cb) {
  const int64_t k = (int64_t)n_scenes * kHitClasses * (1ll << (3 * cb));
  return (int)((k + kHitKeyQuantum - 1) / kHitKeyQuantum * kHitKeyQuantum);
}
// Layout of the sort's buffer for n paths: keys, ranks and order at a stride of
// n rounded up to 64 ints (the counts stay 16-byte aligned for the scan's
// vector loads), then counts, starts and the total.
inline size_t hit_sort_stride(size_t n) { return (n + 63) & ~(size_t)63; }
inline size_t hit_sort_ints(size_t n) { return 3 * hit_sort_stride(n) + 2 * (size_t)kHitMaxKeys + 64; }
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // bits 0-9 to bits 0,3,..,27
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ int hit_key(const DevMesh& m, F3 p, int cb) {
  const F3 pl = xf_to_local(m.xf, p);
  float lo, hi;
  mesh_dist_bounds_local(m, pl, lo, hi);
  const float mx = fmaxf(m.lbmax[0] - m.lbmin[0], fmaxf(m.lbmax[1] - m.lbmin[1], m.lbmax[2] - m.lbmin[2]));
  const float r = hi / mx;
  const int b = r < 0.02f ? 0 : (r < 0.06f ? 1 : (r < 0.15f ? 2 : 3));
  const float pad = 0.25f * mx;
  const float q[3] = {pl.x, pl.y, pl.z};
  uint32_t morton = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float lo_k = m.lbmin[k] - pad, ext = m.lbmax[k] - m.lbmin[k] + 2.0f * pad;
    const float u = (q[k] - lo_k) / fmaxf(ext, 1e-30f) * (float)(1 << cb);
    // NaN and far points land in range: fminf/fmaxf return the number operand
    const int c = (int)fminf(fmaxf(u, 0.0f), (float)((1 << cb) - 1));
    morton |= spread3((uint32_t)c) << k;
  }
  return (b << (3 * cb)) | (int)morton;
}
__global__ __launch_bounds__(256) void k_hit_zero(int* hist, int nkeys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nkeys) hist[i] = 0;
}
// Entry i of the concatenation of queues a (na entries) and b (nb).
__device__ __forceinline__ int hit_queue_at(const int* qa, int na, const int* qb, int i) {
  return i < na ? qa[i] : qb[i - na];
}
// kMulti: a batched launch, each path's mesh from its scene; paths of scenes
// without a GP medium with a mesh mean get no key (the shading kernel reads no
// mean for them).
template <bool kMulti>
__global__ __launch_bounds__(256) void k_hit_key(DevMesh m0, const DevScene* scenes,
                                                 const uint16_t* sid, const int* qa,
                                                 const int* na, const int* qb, const int* nb,
                                                 const uint8_t* mhit, const float* px,
                                                 const float* py, const float* pz, int cb,
                                                 int* keys, int* ranks, int* hist) {
  const int n0 = *na, n = n0 + (nb ? *nb : 0);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (__ballot(i < n) == 0) return;  // whole waves past the queue
  int key = -1;
  if (i < n) {
    const int pid = hit_queue_at(qa, n0, qb, i);
    if (mhit[pid] == 1) {
      const F3 p = mk(px[pid], py[pid], pz[pid]);
      if constexpr (kMulti) {
        const int s = sid[pid];
        const DevScene& S = scenes[s];
        if (S.medium_kind == MUPSI_MEDIUM_GP && S.mean_type == MUPSI_MEAN_MESH &&
            S.mean_mesh.n_nodes > 0)
          key = s * (kHitClasses << (3 * cb)) + hit_key(S.mean_mesh, p, cb);
      } else {
        key = hit_key(m0, p, cb);
      }
    }
  }
  // one atomic per distinct key of the wave: the leader of each key adds the
  // key's lane count and hands the old value to its lanes
  const int lane = __lane_id();
  uint64_t todo = __ballot(key >= 0);
  int rank = -1;
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int lk = __shfl(key, leader);
    const uint64_t mm = __ballot(key == lk);
    int base = 0;
    if (lane == leader) base = atomicAdd(hist + lk, __popcll(mm));
    base = __shfl(base, leader);
    if (key == lk) rank = base + __popcll(mm & ((1ull << lane) - 1ull));
    todo &= ~mm;
  }
  if (i < n) {
    keys[i] = key;
    ranks[i] = rank;
  }
}
// One block: hist (nkeys counts, a multiple of kHitKeyQuantum) -> start
// (exclusive prefix sums); the total goes to *total.
__global__ __launch_bounds__(kHitScanThreads) void k_hit_scan(const int* hist, int* start,
                                                              int* total, int nkeys) {
  __shared__ int part[kHitScanThreads];
  const int t = threadIdx.x;
  const int per4 = nkeys / kHitKeyQuantum;  // int4 loads per thread
  const int4* h4 = reinterpret_cast<const int4*>(hist) + (size_t)t * per4;
  int sum = 0;
  for (int k = 0; k < per4; ++k) {
    const int4 v = h4[k];
    sum += v.x + v.y + v.z + v.w;
  }
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < kHitScanThreads; o <<= 1) {  // inclusive Hillis-Steele scan
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  int4* s4 = reinterpret_cast<int4*>(start) + (size_t)t * per4;
  for (int k = 0; k < per4; ++k) {
    const int4 v = h4[k];
    int4 o;
    o.x = run;
    o.y = o.x + v.x;
    o.z = o.y + v.y;
    o.w = o.z + v.z;
    run = o.w + v.w;
    s4[k] = o;
  }
  if (t == kHitScanThreads - 1) *total = part[t];
}
__global__ __launch_bounds__(256) void k_hit_place(const int* qa, const int* na, const int* qb,
                                                   const int* nb, const int* keys,
                                                   const int* ranks, const int* start,
                                                   int* order) {
  const int n0 = *na, n = n0 + (nb ? *nb : 0);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int key = keys[i];
  if (key >= 0) order[start[key] + ranks[i]] = hit_queue_at(qa, n0, qb, i);
}
// The mesh mean and its gradient at every medium hit of a bounce iteration (the
// march queues' paths with a hit), in key order, for the shading kernel: the
// closest-point search at a hit point is a chain of dependent node loads, which
// inside the shading kernel ran on its few hit lanes at that kernel's occupancy
// and without the LDS copy of the tree's top.  Same function, same bits.
// Single-scene launches copy the tree's top to LDS; batched ones take each
// path's mesh from its scene (no LDS copy: the paths' meshes differ).
template <bool kMulti>
__global__ __launch_bounds__(256) void k_hit_mean_sorted(DevMesh m0, const DevScene* scenes,
                                                         const uint16_t* sid, const int* order,
                                                         const int* total, const float* px,
                                                         const float* py, const float* pz,
                                                         float4* out) {
  __shared__ BvhRec top[kMulti ? 1 : kTopRecs];
  const int n = *total;
  if (n <= 0) return;  // the whole grid: no hit to serve
  DevMesh m;
  if constexpr (!kMulti) m = stage_top(m0, top);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int pid = order[j];
    F3 g;
    const F3 p = mk(px[pid], py[pid], pz[pid]);
    float mu;
    if constexpr (kMulti) mu = mesh_mean_eval<true, false>(scenes[sid[pid]].mean_mesh, p, &g);
    else mu = mesh_mean_eval<true, true>(m, p, &g);
    out[pid] = make_float4(mu, g.x, g.y, g.z);
  }
}
// The sort and the searches for the hits of queues a and b (b may be null).
inline void launch_hit_means(hipStream_t st, int n, int query_blocks, bool multi, const DevMesh& m,
                             const DevScene* scenes, const uint16_t* sid, int n_scenes,
                             const int* qa, const int* na, const int* qb, const int* nb,
                             const PathSoA& P, int* buf, size_t capacity, float4* out) {
  const size_t cap = hit_sort_stride(capacity);
  // a single mesh of fewer than 65,536 triangles: distance classes only (its
  // tree is small enough that the cells gained less than the larger scan cost)
  const int cb = multi ? hit_cell_bits(n_scenes) : (m.n_faces >= 65536 ? kHitCellBits : 0);
  const int nkeys = hit_key_count(multi ? n_scenes : 1, cb);
  int* keys = buf;
  int* ranks = keys + cap;
  int* order = ranks + cap;
  int* hist = order + cap;
  int* start = hist + kHitMaxKeys;
  int* total = start + kHitMaxKeys;
  const dim3 g(grid_for(n));
  hipLaunchKernelGGL(k_hit_zero, dim3(nkeys / 256), dim3(256), 0, st, hist, nkeys);
  if (multi)
    hipLaunchKernelGGL(k_hit_key<true>, g, dim3(256), 0, st, DevMesh{}, scenes, sid, qa, na, qb,
                       nb, P.mhit, P.mpx, P.mpy, P.mpz, cb, keys, ranks, hist);
  else
    hipLaunchKernelGGL(k_hit_key<false>, g, dim3(256), 0, st, m, nullptr, nullptr, qa, na, qb, nb,
                       P.mhit, P.mpx, P.mpy, P.mpz, cb, keys, ranks, hist);
  hipLaunchKernelGGL(k_hit_scan, dim3(1), dim3(kHitScanThreads), 0, st, hist, start, total, nkeys);
  hipLaunchKernelGGL(k_hit_place, g, dim3(256), 0, st, qa, na, qb, nb, keys, ranks, start, order);
  const dim3 hg(std::min<unsigned>(g.x, (unsigned)std::max(1, query_blocks)));
  if (multi)
    hipLaunchKernelGGL(k_hit_mean_sorted<true>, hg, dim3(256), 0, st, DevMesh{}, scenes, sid,
                       order, total, P.mpx, P.mpy, P.mpz, out);
  else
    hipLaunchKernelGGL(k_hit_mean_sorted<false>, hg, dim3(256), 0, st, m, nullptr, nullptr, order,
                       total, P.mpx, P.mpy, P.mpz, out);
}

// Traversal cost per point of Mesh::distToPoint (kind 0), Mesh::inside (kind 1), or
// Mesh::distToPoint searched from the distance grid's bound (kind 2), or inside
// as mupsi_mesh_inside runs it (kind 3), or inside as the stackless walk over the
// tree (kind 4; kind 3 where the mesh has no leaf bins).
__global__ __launch_bounds__(256) void k_mesh_cost(DevMesh m, int kind, const float* px,
                                                   const float* py, const float* pz, int64_t n,
                                                   uint32_t* nodes, uint32_t* tris) {
  MESH_QUERY_LOOP(i, n) {
    TravCount c;
    const F3 p = mk(px[i], py[i], pz[i]);
    if (kind == 0 || kind == 2) {
      uint32_t ti = 0;
      F3 o = mk(0, 0, 0);
      if (kind == 0) (void)mesh_dist(m, p, ti, o, c);
      else (void)mesh_dist_grid(m, p, ti, o, c);
    } else if (kind == 1) {
      (void)mesh_inside<0>(m, p, c);  // the reference's casts
    } else if (kind == 4) {
      (void)mesh_inside<1>(m, p, c);  // the stackless walk without the leaf bins
    } else {
      (void)mesh_inside<2>(m, p, c);  // as the library runs it
    }
    nodes[i] = c.nodes;
    tris[i] = c.tris;
  }
}

// ====================================================================== no-allocation scratch

// Device scratch of mupsi_march_noalloc, a file-scope array instead of per-call
// allocations: a 256-byte head (the persistent marcher's work counter at byte 0,
// its four 64-bit statistics at byte 8) and nine arrays of kNoAllocMaxRays
// floats (normalized ray origin and direction, hit point).  Each GPU the library
// runs on holds its own copy until the process ends.  Host code never passes
// the name `scratch` to a kernel (on the host it is a stand-in, not a device
// address): it uses the per-device address from noalloc_scratch().
constexpr int64_t kNoAllocMaxRays = MUPSI_MARCH_NOALLOC_MAX_RAYS;
constexpr size_t kScratchHead = 256;
constexpr size_t kScratchBytes = kScratchHead + 9 * (size_t)kNoAllocMaxRays * sizeof(float);
__device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_addr;  // device address of `scratch`, per device index

// Looks the array's address up once per device (the caller has made `device`
// current) and keeps it; later calls return the kept address.
static mupsi_status noalloc_scratch(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= device) g_scratch_addr.resize(device + 1, nullptr);
  if (!g_scratch_addr[device]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(MUPSI_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
    g_scratch_addr[device] = ptr;
  }
  *out = g_scratch_addr[device];
  return MUPSI_OK;
}

// ====================================================================== host: context

extern "C" const char* mupsi_status_string(mupsi_status s) {
  switch (s) {
    case MUPSI_OK: return "ok";
    case MUPSI_ERR_INVALID: return "invalid argument";
    case MUPSI_ERR_HIP: return "HIP runtime error";
    case MUPSI_ERR_NOMEM: return "out of memory";
    case MUPSI_ERR_ABORTED: return "aborted";
    case MUPSI_ERR_UNSUPPORTED: return "unsupported";
    case MUPSI_ERR_IO: return "I/O error";
    case MUPSI_ERR_PARSE: return "parse error";
    default: return "unknown status";
  }
}
extern "C" const char* mupsi_last_error(void) { return g_last_error.c_str(); }

extern "C" mupsi_status mupsi_open(int32_t device, mupsi_ctx** out) {
  if (!out) return fail(MUPSI_ERR_INVALID, "mupsi_open: out is null");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(MUPSI_ERR_HIP, "mupsi_open: no HIP device");
  if (device < 0 || device >= count) return fail(MUPSI_ERR_INVALID, "mupsi_open: bad device index");
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(MUPSI_ERR_HIP, std::string("mupsi_open: built for gfx950, device is ") + prop.gcnArchName);
  auto ctx = std::make_unique<mupsi_ctx>();
  ctx->device = device;
  ctx->name = prop.name;
  ctx->cus = prop.multiProcessorCount;
  HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&ctx->ev_begin));
  HIP_TRY(hipEventCreate(&ctx->ev_end));
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_march<false>, 256, 0));
  ctx->march_blocks = std::max(1, per_cu) * ctx->cus;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_march_sphere<3, sphere_group(3)>, 64, 0));
  ctx->march_cached_blocks = std::max(1, per_cu) * ctx->cus;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mesh_closest, 256, 0));
  // at most four per CU: five fit the LDS but ran the 4M-query closest-point
  // search at 46 ms against 34 ms with four (three: 38 ms, two: 50 ms)
  ctx->query_blocks = std::max(1, std::min(per_cu, 4)) * ctx->cus;
  {
    const mupsi_status st = noalloc_scratch(device, &ctx->scratch);
    if (st != MUPSI_OK) return st;
  }
  *out = ctx.release();
  return MUPSI_OK;
}

static void destroy_scene(mupsi_scene* scene);
static void destroy_batch(mupsi_batch* b);
static void batch_forget_scene(mupsi_ctx* ctx, const mupsi_scene* scene);
static bool scene_in_live_batch(const mupsi_ctx* ctx, const mupsi_scene* scene);

// Frees the context and every scene still open on it (their handles become invalid).
extern "C" void mupsi_close(mupsi_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  std::vector<mupsi_batch*> batches;
  batches.swap(ctx->batches);
  for (mupsi_batch* b : batches) destroy_batch(b);
  std::vector<mupsi_scene*> live;
  live.swap(ctx->scenes);
  for (mupsi_scene* s : live) destroy_scene(s);
  std::vector<mupsi_guide*> guides;
  guides.swap(ctx->guides);
  for (mupsi_guide* g : guides) destroy_guide(g);
  if (ctx->ev_begin) (void)hipEventDestroy(ctx->ev_begin);
  if (ctx->ev_end) (void)hipEventDestroy(ctx->ev_end);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" mupsi_status mupsi_device_info(mupsi_ctx* ctx, char* name, int32_t name_len,
                                          int32_t* cu_count) {
  if (!ctx) return fail(MUPSI_ERR_INVALID, "mupsi_device_info: null ctx");
  if (name && name_len > 0) {
    std::strncpy(name, ctx->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = ctx->cus;
  return MUPSI_OK;
}

// ====================================================================== host: scene

struct PathBuffers {
  DeviceBuffers mem;
  int64_t capacity = 0;
  PathSoA P{};
  int* qA[2]{};
  int* qM[2]{};
  int* counters = nullptr;  // [nA0, nM0, nA1, nM1, work, nG, training records, nS,
                            //  batched: nMm0, nMm1, work of queue Mm]
  unsigned long long* stats = nullptr;
  ShadowSoA SH{};
  int* qS = nullptr;
  int* qMm[2]{};  // batched launches: medium paths for the lane-per-ray marcher
  int* qMs[2]{};  // batched launches: window-marcher paths of mesh-mean scenes
  float4* hmean = nullptr;  // k_hit_mean's output, one per path
  int* hbins = nullptr;     // single-scene launches: keys, ranks and order of the hits
                            // (capacity each), then the key counts, starts and total
};

// Buffers of guided rendering: per-path mixture and features of the current
// bounce, queue G, and the training records of one pass.
struct GuideBuffers {
  DeviceBuffers mem;
  int64_t capacity = 0;      // paths
  int64_t rec_capacity = 0;  // record slots (paths of a training pass x K)
  float* mix = nullptr;      // capacity x 44
  float* feat = nullptr;     // capacity x 32
  int* qG = nullptr;
  float* rec = nullptr;      // rec_capacity x 10
  float* rec_feat = nullptr; // rec_capacity x 32
  // training samples: a ring of tr_capacity records kept across the training
  // passes (minibatches are drawn from all of it)
  int64_t tr_capacity = 0, tr_head = 0, tr_size = 0;
  float* tr_feat = nullptr;  // tr_capacity x 32
  float* tr_dir = nullptr;   // tr_capacity x 3
  float* tr_w = nullptr;     // tr_capacity
};

using EventPairs = std::vector<std::pair<hipEvent_t, hipEvent_t>>;

static void destroy_events(EventPairs& evs) {
  for (auto& e : evs) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  evs.clear();
}
static double elapsed_ms(const EventPairs& evs) {
  double ms = 0;
  for (const auto& e : evs) {
    float t = 0;
    (void)hipEventElapsedTime(&t, e.first, e.second);
    ms += t;
  }
  return ms;
}

// A captured plain-tracer wavefront of n paths over the pixel grid: the graph
// reads its first pass from kb_dev, so one capture serves every later launch
// of the same size.
struct WavefrontGraph {
  hipGraphExec_t exec = nullptr;
  int64_t n = -1;
  int* kb_dev = nullptr;
  EventPairs evs, gevs;  // empty: a capture records no events
  DeviceBuffers mem;
  ~WavefrontGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

struct mupsi_scene {
  mupsi_ctx* ctx = nullptr;
  DevScene ds{};
  DeviceBuffers mem;
  HostScene host;  // host-side lowered description (meshes keep the BVH for the batch API)
  std::vector<DevMesh> meshes;
  DevMesh mean_mesh{};
  std::unique_ptr<DeviceBuffers> mean_mem;  // mean mesh + distance grid (replaceable)
  double* accum = nullptr;
  int64_t sample_count = 0;
  std::unique_ptr<PathBuffers> pb;
  std::unique_ptr<GuideBuffers> gb;
  mupsi_guide* guide = nullptr;  // attached guide, or null (plain tracer)
  mupsi_stats last{};
  // plain renders replay a captured wavefront (mupsi_scene_set_graph); the
  // capture holds pointers to pb and the scene constants, so it is dropped
  // whenever either changes
  int use_graph = 0;
  int64_t graph_builds = 0;
  std::unique_ptr<WavefrontGraph> graph;
  // the marcher shades GP-medium hits itself where the scene allows it
  // (mupsi_scene_set_fused; fuse_march)
  int fused = 1;
};

// The scene holds a triangle mesh, as a primitive or as the mean: the shading
// kernel then needs the BVH code (its sphere-only instance has none).
static bool scene_has_tri(const mupsi_scene* sc) {
  return !sc->meshes.empty() || sc->ds.mean_type == MUPSI_MEAN_MESH;
}

// Distances from the centres of a grid of cubes to a mean mesh (local space): for
// any point of a cube the distance is within the half diagonal of the centre's
// (distance is 1-Lipschitz).
// Distance from the centre of every cube of a grid (origin org, n cubes per axis
// of side cell) to the mesh, in local space.
__global__ __launch_bounds__(256) void k_dist_grid(DevMesh m0, float ox, float oy, float oz, int nx,
                                                   int ny, int nz, float cell, float* out) {
  __shared__ BvhRec top[kTopRecs];
  const DevMesh m = stage_top(m0, top);
  const int n = nx * ny * nz;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
  const F3 c = mk(ox + (ix + 0.5f) * cell, oy + (iy + 0.5f) * cell, oz + (iz + 0.5f) * cell);
  uint32_t ti = 0;
  F3 o;
  const float d = bvh_closest<true>(m, c, ti, o);
  out[i] = d;
}

// One distance grid of k cubes along the longest axis over the mesh's local box
// padded by pad_rel times that axis's length on every side.
static mupsi_status make_dist_grid(DeviceBuffers& mem, const DevMesh& m, int k, float pad_rel,
                                   const float** g, int* n, float* org, float* inv, float* cellp,
                                   float* slack) {
  float mx = 0.0f;
  for (int a = 0; a < 3; ++a) mx = std::max(mx, m.lbmax[a] - m.lbmin[a]);
  const float pad = pad_rel * mx;
  const float cell = (mx + 2.0f * pad) / k;
  for (int a = 0; a < 3; ++a) {
    org[a] = m.lbmin[a] - pad;
    n[a] = std::max(1, std::min(k, (int)std::ceil((m.lbmax[a] - m.lbmin[a] + 2.0f * pad) / cell)));
  }
  *inv = 1.0f / cell;
  *cellp = cell;
  // rounding margin of the centre distance and of |p - c|: a few ulps of the
  // largest coordinate in the grid
  *slack = 1e-6f * (1.0f + mx + 2.0f * pad);
  const int cnt = n[0] * n[1] * n[2];
  float* grid = nullptr;
  HIP_TRY(mem.alloc(cnt, &grid));
  DevMesh mm = m;
  mm.dlb = nullptr;
  mm.olb = nullptr;
  mm.flb = nullptr;
  hipLaunchKernelGGL(k_dist_grid, dim3(grid_for(cnt)), dim3(256), 0, 0, mm, org[0], org[1], org[2],
                     n[0], n[1], n[2], cell, grid);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  *g = grid;
  return MUPSI_OK;
}

static mupsi_status build_dist_grid(DeviceBuffers& mem, DevMesh* m) {
  float mx = 0.0f;
  for (int a = 0; a < 3; ++a) mx = std::max(mx, m->lbmax[a] - m->lbmin[a]);
  if (m->n_nodes == 0 || !(mx > 0.0f) || !std::isfinite(mx)) return MUPSI_OK;
  // 128 cubes along the longest axis over the box padded by a quarter (256 for
  // meshes of 65,536 triangles or more, whose exact searches are the longest): an
  // evaluation needs the exact distance only where the bound straddles the
  // surface, a band about |p - c| wide
  const int fine = m->n_faces >= 65536 ? 256 : 128;
  mupsi_status st = make_dist_grid(mem, *m, fine, 0.25f, &m->dlb, m->dn, m->dorg, &m->dinv,
                                   &m->dcell, &m->dslack);
  if (st != MUPSI_OK) return st;
  // 96 cubes over the box padded by 1.5 times its length: points farther out
  // are evaluated where the field's noise and conditioning terms can reach the
  // distance there, and each such point would take an unbounded exact search
  st = make_dist_grid(mem, *m, 96, 1.5f, &m->olb, m->on, m->oorg, &m->oinv, &m->ocell,
                      &m->oslack);
  if (st != MUPSI_OK) return st;
  // 64 cubes over the box padded by 8 times its length: Renewal+ multiplies the
  // conditioning by about 1/sigma at every medium hit, so after many bounces of
  // a low-sigma medium it reaches the mean's value far from the mesh
  return make_dist_grid(mem, *m, 64, 8.0f, &m->flb, m->fn, m->forg, &m->finv, &m->fcell,
                        &m->fslack);
}

// Skip links of a left-first pre-order walk (the next node once a node's subtree
// is done or skipped; -1 ends the walk), if the tree allows the stackless inside
// walk: every node reachable once, indices in range, and a depth at which the
// reference's near/far traversal (at most one waiting sibling per level) can
// never fill its 64-entry stack, so no cast of the reference drops a subtree.
// with_bins: also the leaf bins of inside() (mean meshes, whose inside() the
// marcher queries at every grid point; trees that allow the stackless walk)
static mupsi_status upload_mesh(DeviceBuffers& mem, const HostMesh& hm, DevMesh* out,
                                bool with_bins = false) {
  DevMesh m{};
  BvhRecords L;
  if (!build_records(hm.bvh, hm.order, L))
    return fail(MUPSI_ERR_INVALID, "mesh upload: the BVH does not fit the device record layout");
  const BvhRec* recs;
  HIP_TRY(mem.upload(L.recs.data(), L.recs.size(), &recs));
  m.recs = recs;
  m.top = nullptr;
  m.ntop = 0;
  m.n_recs = (uint32_t)L.recs.size();
  m.root_ref = L.root_ref;
  m.occ_ref = L.occ_ref;
  for (int a = 0; a < 3; ++a) {
    m.occ_min[a] = L.occ_min[a];
    m.occ_max[a] = L.occ_max[a];
  }
  m.stackless = L.depth <= dev::kStack - 4 ? 1 : 0;
  HIP_TRY(mem.upload(hm.leaf_tris.data(), hm.leaf_tris.size(), &m.leaf_tris));
  HIP_TRY(mem.upload(hm.vnormals.data(), hm.vnormals.size(), &m.vnormals));
  HIP_TRY(mem.upload(hm.faces.data(), hm.faces.size(), &m.faces));
  HIP_TRY(mem.upload(hm.fnormals.data(), hm.fnormals.size(), &m.fnormals));
  HIP_TRY(mem.upload(hm.texcoords.data(), hm.texcoords.size(), &m.texcoords));
  HIP_TRY(mem.upload(hm.face_tex.data(), hm.face_tex.size(), &m.face_tex));
  HIP_TRY(mem.upload(hm.verts.data(), hm.verts.size(), &m.verts));
  HIP_TRY(mem.upload(hm.face_cdf.data(), hm.face_cdf.size(), &m.face_cdf));
  m.total_area = hm.total_area;
  m.n_nodes = (int)hm.bvh.nodes.size();
  m.n_faces = (int)(hm.faces.size() / 3);
  m.xf = hm.xf;
  for (int a = 0; a < 3; ++a) {
    m.lbmin[a] = hm.bvh.nodes.empty() ? 0.f : hm.bvh.nodes[hm.bvh.root].bmin[a];
    m.lbmax[a] = hm.bvh.nodes.empty() ? 0.f : hm.bvh.nodes[hm.bvh.root].bmax[a];
  }
  XBins xb;
  if (with_bins && m.stackless && build_xbins(L, m.lbmin, m.lbmax, xb)) {
    const XBinEntry* ent;
    HIP_TRY(mem.upload(xb.entries.data(), xb.entries.size(), &ent));
    HIP_TRY(mem.upload(xb.offsets.data(), xb.offsets.size(), &m.xb_off));
    m.xb_ent = reinterpret_cast<const float4*>(ent);
    m.xb_n[0] = xb.ny;
    m.xb_n[1] = xb.nz;
    m.xb_org[0] = xb.y0;
    m.xb_org[1] = xb.z0;
    m.xb_inv = xb.inv;
  }
  *out = m;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_scene_upload(mupsi_ctx* ctx, const mupsi_scene_desc* desc,
                                           mupsi_scene** out) {
  if (!ctx || !desc || !out) return fail(MUPSI_ERR_INVALID, "mupsi_scene_upload: null argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  auto sc = std::make_unique<mupsi_scene>();
  sc->ctx = ctx;
  std::string err;
  mupsi_status st = lower_scene(*desc, &sc->host, &err);
  if (st != MUPSI_OK) return fail(st, err);
  const HostScene& H = sc->host;
  DevScene& D = sc->ds;
  D = H.ds;  // constants computed on the host

  for (const HostMesh& hm : H.meshes) {
    DevMesh m;
    st = upload_mesh(sc->mem, hm, &m);
    if (st != MUPSI_OK) return st;
    sc->meshes.push_back(m);
  }
  sc->mean_mem = std::make_unique<DeviceBuffers>();
  if (H.has_mean_mesh) {
    st = upload_mesh(*sc->mean_mem, H.mean_mesh, &sc->mean_mesh, true);
    if (st != MUPSI_OK) return st;
    st = build_dist_grid(*sc->mean_mem, &sc->mean_mesh);
    if (st != MUPSI_OK) return st;
  }
  D.mean_mesh = sc->mean_mesh;
  const DevMesh* dmeshes;
  HIP_TRY(sc->mem.upload(sc->meshes.data(), sc->meshes.size(), &dmeshes));
  D.meshes = dmeshes;

  std::vector<DevPrim> prims = H.prims;
  for (size_t i = 0; i < prims.size(); ++i) {
    const std::vector<float>& tex = H.prim_albedo[i];
    const float* dtex = nullptr;
    HIP_TRY(sc->mem.upload(tex.data(), tex.size(), &dtex));
    prims[i].albedo.rgb = dtex;
  }
  const DevPrim* dprims;
  HIP_TRY(sc->mem.upload(prims.data(), prims.size(), &dprims));
  D.prims = dprims;
  const float* dsky = nullptr;
  HIP_TRY(sc->mem.upload(H.sky.data(), H.sky.size(), &dsky));
  D.sky.rgb = dsky;
  const float *dgm = nullptr, *dgv = nullptr;
  HIP_TRY(sc->mem.upload(H.grid_mean.data(), H.grid_mean.size(), &dgm));
  HIP_TRY(sc->mem.upload(H.grid_var.data(), H.grid_var.size(), &dgv));
  D.grid.mean = dgm;
  D.grid.var = dgv;

  const size_t npix = (size_t)D.w * D.h;
  HIP_TRY(sc->mem.alloc(npix * 3, &sc->accum));
  HIP_TRY(hipMemset(sc->accum, 0, npix * 3 * sizeof(double)));
  ctx->scenes.push_back(sc.get());
  *out = sc.release();
  return MUPSI_OK;
}

static void destroy_scene(mupsi_scene* scene) {
  (void)hipStreamSynchronize(scene->ctx->stream);
  delete scene;
}

extern "C" void mupsi_scene_free(mupsi_scene* scene) {
  if (!scene) return;
  mupsi_ctx* ctx = scene->ctx;
  auto it = std::find(ctx->scenes.begin(), ctx->scenes.end(), scene);
  if (it != ctx->scenes.end()) ctx->scenes.erase(it);
  batch_forget_scene(ctx, scene);
  (void)hipSetDevice(ctx->device);
  destroy_scene(scene);
}

// Controller::applyGpMedium (controller.cpp:397-426) on a live scene: a new
// GPMedium for every int_medium primitive.  The sums are kept (the caller resets
// them as Renderer::prepareRender does); batches hold their own copy of the
// scene, so a scene in a live batch is refused.
extern "C" mupsi_status mupsi_scene_set_gp_medium(mupsi_scene* scene, int32_t mean_type,
                                                  const float* mean_center, float mean_radius,
                                                  const mupsi_mesh_desc* mean_mesh,
                                                  const mupsi_gp_params* kernel) {
  if (!scene || !kernel) return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_gp_medium: null argument");
  if (scene_in_live_batch(scene->ctx, scene))
    return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_gp_medium: scene belongs to a live batch");
  HIP_TRY(hipSetDevice(scene->ctx->device));
  HIP_TRY(hipStreamSynchronize(scene->ctx->stream));  // no launch still reads the old mean
  HostScene& H = scene->host;
  const bool had_mesh = H.has_mean_mesh;
  HostMesh old_mesh = std::move(H.mean_mesh);
  const DevScene old_ds = H.ds;
  std::string err;
  mupsi_status st = lower_gp_medium(mean_type, mean_center, mean_radius, mean_mesh, *kernel, &H, &err);
  if (st != MUPSI_OK) {
    H.mean_mesh = std::move(old_mesh);
    return fail(st, err);
  }
  auto mem = std::make_unique<DeviceBuffers>();
  DevMesh mm{};
  if (H.has_mean_mesh) {
    st = upload_mesh(*mem, H.mean_mesh, &mm, true);
    if (st == MUPSI_OK) st = build_dist_grid(*mem, &mm);
    if (st != MUPSI_OK) {
      H.ds = old_ds;
      H.has_mean_mesh = had_mesh;
      H.mean_mesh = std::move(old_mesh);
      return st;
    }
  }
  scene->graph.reset();  // the capture holds the old scene constants
  DevScene& D = scene->ds;
  D.gp = H.ds.gp;
  D.mean_type = H.ds.mean_type;
  for (int a = 0; a < 3; ++a) D.mean_c[a] = H.ds.mean_c[a];
  D.mean_r = H.ds.mean_r;
  scene->mean_mesh = mm;
  D.mean_mesh = mm;
  scene->mean_mem = std::move(mem);
  return MUPSI_OK;
}

// The render settings a RenderContext snapshots (render_context.h:15-19,
// trace.h:16-23, gpnoise.h GPSettings): bounce limit, GP correlation mode and
// skip_medium, as Controller::setGpMode and startProgressive change them.
extern "C" mupsi_status mupsi_scene_set_settings(mupsi_scene* scene, int32_t max_bounce,
                                                 int32_t gp_mode, int32_t skip_medium) {
  if (!scene) return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_settings: null scene");
  if (scene_in_live_batch(scene->ctx, scene))
    return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_settings: scene belongs to a live batch");
  if (gp_mode != MUPSI_GP_SINGLE_REALIZATION && gp_mode != MUPSI_GP_RENEWAL_PLUS)
    return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_settings: unknown gp_mode");
  HIP_TRY(hipSetDevice(scene->ctx->device));
  HIP_TRY(hipStreamSynchronize(scene->ctx->stream));
  scene->graph.reset();  // the capture holds the old settings
  for (DevScene* d : {&scene->ds, &scene->host.ds}) {
    d->max_bounce = max_bounce;
    d->gp_mode = gp_mode == MUPSI_GP_RENEWAL_PLUS ? 1 : 0;
    d->skip_medium = skip_medium ? 1 : 0;
  }
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_scene_size(mupsi_scene* scene, int32_t* width, int32_t* height) {
  if (!scene) return fail(MUPSI_ERR_INVALID, "mupsi_scene_size: null scene");
  if (width) *width = scene->ds.w;
  if (height) *height = scene->ds.h;
  return MUPSI_OK;
}

// hit_means: buffers of the mesh-mean hit kernels (the means, and for
// single-scene launches, mesh_queue false, their bins)
static mupsi_status alloc_path_buffers(std::unique_ptr<PathBuffers>& out, int64_t n, bool shadow,
                                       bool mesh_queue, bool hit_means) {
  out.reset();
  auto pb = std::make_unique<PathBuffers>();
  DeviceBuffers& m = pb->mem;
  PathSoA& P = pb->P;
  HIP_TRY(m.alloc(n, &P.rng));
  float** f32[] = {&P.ox, &P.oy, &P.oz, &P.dx, &P.dy, &P.dz, &P.nt, &P.Tx, &P.Ty, &P.Tz, &P.mis,
                   &P.Lx, &P.Ly, &P.Lz, &P.cx, &P.cy, &P.cz, &P.cu, &P.cgx, &P.cgy, &P.cgz,
                   &P.mt, &P.mpx, &P.mpy, &P.mpz};
  for (float** p : f32) HIP_TRY(m.alloc(n, p));
  HIP_TRY(m.alloc(n, &P.flags));
  HIP_TRY(m.alloc(n, &P.mseed));
  HIP_TRY(m.alloc(n, &P.mhit));
  HIP_TRY(m.alloc(n, &P.sid));
  if (shadow) {  // deferred medium light samples
    ShadowSoA& H = pb->SH;
    float** sf[] = {&H.ox, &H.oy, &H.oz, &H.dx, &H.dy, &H.dz, &H.nt, &H.ft, &H.cx, &H.cy, &H.cz,
                    &H.cu, &H.cgx, &H.cgy, &H.cgz, &H.t, &H.px, &H.py, &H.pz, &H.pend_x,
                    &H.pend_y, &H.pend_z, &H.ac};
    for (float** p : sf) HIP_TRY(m.alloc(n, p));
    HIP_TRY(m.alloc(n, &H.seed));
    HIP_TRY(m.alloc(n, &H.active));
    HIP_TRY(m.alloc(n, &H.hit));
    HIP_TRY(m.alloc(n, &pb->qS));
  }
  for (int i = 0; i < 2; ++i) {
    HIP_TRY(m.alloc(n, &pb->qA[i]));
    HIP_TRY(m.alloc(n, &pb->qM[i]));
    if (mesh_queue) {
      HIP_TRY(m.alloc(n, &pb->qMm[i]));
      HIP_TRY(m.alloc(n, &pb->qMs[i]));
    }
  }
  HIP_TRY(m.alloc(16, &pb->counters));
  HIP_TRY(m.alloc(8, &pb->stats));
  if (hit_means) {
    HIP_TRY(m.alloc(n, &pb->hmean));
    HIP_TRY(m.alloc(hit_sort_ints(n), &pb->hbins));
  }
  pb->capacity = n;
  out = std::move(pb);
  return MUPSI_OK;
}

static mupsi_status ensure_paths(mupsi_scene* sc, int64_t n) {
  // a mesh mean (possibly set after the buffers were made) needs the hit buffers
  const bool hit_means = sc->ds.mean_type == MUPSI_MEAN_MESH;
  if (sc->pb && sc->pb->capacity >= n && (!hit_means || sc->pb->hbins)) return MUPSI_OK;
  sc->graph.reset();  // it holds pointers into the old buffers
  return alloc_path_buffers(sc->pb, n, sc->ds.n_lights > 0 && sc->ds.has_gp, false, hit_means);
}

static mupsi_status ensure_guide_buffers(mupsi_scene* sc, int64_t n, int64_t rec_slots,
                                         int64_t ring) {
  GuideBuffers* gb = sc->gb.get();
  if (gb && gb->capacity >= n && gb->rec_capacity >= rec_slots && gb->tr_capacity >= ring)
    return MUPSI_OK;
  const int64_t cap = std::max(n, gb ? gb->capacity : 0);
  const int64_t rcap = std::max(rec_slots, gb ? gb->rec_capacity : 0);
  const int64_t tcap = std::max(ring, gb ? gb->tr_capacity : 0);
  sc->gb.reset();
  auto g = std::make_unique<GuideBuffers>();
  HIP_TRY(g->mem.alloc(cap * guide::kMix, &g->mix));
  HIP_TRY(g->mem.alloc(cap * guide::kIn, &g->feat));
  HIP_TRY(g->mem.alloc(cap, &g->qG));
  HIP_TRY(g->mem.alloc(rcap * kRecFloats, &g->rec));
  HIP_TRY(g->mem.alloc(rcap * guide::kIn, &g->rec_feat));
  HIP_TRY(g->mem.alloc(tcap * guide::kIn, &g->tr_feat));
  HIP_TRY(g->mem.alloc(tcap * 3, &g->tr_dir));
  HIP_TRY(g->mem.alloc(tcap, &g->tr_w));
  g->capacity = cap;
  g->rec_capacity = rcap;
  g->tr_capacity = tcap;
  sc->gb = std::move(g);
  return MUPSI_OK;
}

void detach_guide(mupsi_scene* s, const mupsi_guide* g) {
  if (s && s->guide == g) s->guide = nullptr;
}

extern "C" mupsi_status mupsi_scene_set_graph(mupsi_scene* sc, int32_t use_graph) {
  if (!sc) return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_graph: null scene");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  HIP_TRY(hipStreamSynchronize(sc->ctx->stream));
  sc->use_graph = use_graph ? 1 : 0;
  if (!sc->use_graph) sc->graph.reset();
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_scene_set_fused(mupsi_scene* sc, int32_t fused) {
  if (!sc) return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_fused: null scene");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  HIP_TRY(hipStreamSynchronize(sc->ctx->stream));
  sc->fused = fused ? 1 : 0;
  sc->graph.reset();  // the capture holds the march arguments
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_scene_graph_builds(mupsi_scene* sc, int64_t* builds) {
  if (!sc || !builds) return fail(MUPSI_ERR_INVALID, "mupsi_scene_graph_builds: null argument");
  *builds = sc->graph_builds;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_scene_set_guide(mupsi_scene* sc, mupsi_guide* g) {
  if (!sc) return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_guide: null scene");
  if (g && g->ctx != sc->ctx) return fail(MUPSI_ERR_INVALID, "mupsi_scene_set_guide: guide of another context");
  sc->guide = g;
  if (g) g->stats.train_passes_done = 0;
  if (sc->gb) sc->gb->tr_head = sc->gb->tr_size = 0;  // a fresh training history
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_scene_guide_records(mupsi_scene* sc, float* feats, float* dirs,
                                                  float* weights, int64_t capacity, int64_t* count) {
  if (!sc || !count) return fail(MUPSI_ERR_INVALID, "mupsi_scene_guide_records: null argument");
  *count = 0;
  if (!sc->gb || sc->gb->tr_size == 0) return MUPSI_OK;
  const int64_t n = std::min<int64_t>(capacity, sc->gb->tr_size);
  if (n <= 0) return MUPSI_OK;
  if (!feats || !dirs || !weights) return fail(MUPSI_ERR_INVALID, "mupsi_scene_guide_records: null array");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  HIP_TRY(hipStreamSynchronize(sc->ctx->stream));
  HIP_TRY(hipMemcpy(feats, sc->gb->tr_feat, n * guide::kIn * sizeof(float), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(dirs, sc->gb->tr_dir, n * 3 * sizeof(float), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(weights, sc->gb->tr_w, n * sizeof(float), hipMemcpyDeviceToHost));
  *count = n;
  return MUPSI_OK;
}

// The cached-window marcher serves every scene with points_per_cell <= 8 (mesh
// means evaluate their BVH only near the surface, see cached_field); larger ppc
// uses the lane-per-ray k_march.
static CachedConsts cached_consts(const DevGP& g) {
  CachedConsts cc;
  cc.inv_kr = 1.0f / g.kr;
  int e = 0;
  cc.kr_pow2 = std::frexp(g.kr, &e) == 0.5f ? 1 : 0;
  cc.psi_k = (float)(-1.4426950408889634 / (double)g.L2);
  cc.psi_scale = (float)((double)g.sigma / (double)g.sqrt_var);
  return cc;
}

static void launch_march(mupsi_ctx* ctx, const DevScene& S, const MarchArgs& ma, hipStream_t st) {
  if (S.medium_kind == MUPSI_MEDIUM_GP && S.gp.ppc <= kMaxPpcCached) {
    const CachedConsts cc = cached_consts(S.gp);
    const dim3 grid(ctx->march_cached_blocks), block(64);
    const bool mesh = S.mean_type == MUPSI_MEAN_MESH;
#define MUPSI_CACHED(P)                                                                     \
  case P:                                                                                   \
    if (mesh)                                                                               \
      hipLaunchKernelGGL((k_march_window<P, true, false, mesh_group(P)>), grid, block, 0, st, \
                         ma, S, cc, nullptr);                                               \
    else                                                                                    \
      hipLaunchKernelGGL((k_march_sphere<P, sphere_group(P)>), grid, block, 0, st, ma, S, cc, \
                         nullptr);                                                          \
    break;
    switch (S.gp.ppc) {
      MUPSI_CACHED(1)
      MUPSI_CACHED(2)
      MUPSI_CACHED(3)
      MUPSI_CACHED(4)
      MUPSI_CACHED(5)
      MUPSI_CACHED(6)
      MUPSI_CACHED(7)
      default: MUPSI_CACHED(8)
    }
#undef MUPSI_CACHED
  } else {
    hipLaunchKernelGGL(k_march<false>, dim3(ctx->march_blocks), dim3(256), 0, st, S, ma);
  }
}

// Run the wavefront over n paths described by src; leaves per-path radiance in P.L*.
struct GuideRun {
  mupsi_guide* g;
  int record;
  double ms;  // guide prep + inference time (HIP events)
};

// Enqueues the wavefront over n paths described by src (init, then the bounce
// loop's marches and shades) on the context stream, leaving per-path radiance
// in P.L*.  With timed, HIP events bracket every march launch (evs) and guide
// inference (gevs); the caller reads them once the stream has drained (a
// graph capture records none).  abort_flag is polled between bounce
// iterations (null: never, as under graph capture).
// The window marcher shades the GP-medium hits of a render itself (handleVolume
// and Renewal+ inside the march launch, MarchArgs::fuse) when the hit needs
// nothing else: a sphere-mean GP medium within the window marcher's points per
// cell, no emitter (no light sample, hence no transmittance march) and no guide.
// (A mesh mean's hit needs the BVH queries of its mean gradient, which would
// hold the whole wave of marching rays while one group traverses: 2.5x slower
// on the 327,680-triangle mean.)
static bool fuse_march(const mupsi_scene* sc, bool guided) {
  const DevScene& S = sc->ds;
  return sc->fused && !guided && S.medium_kind == MUPSI_MEDIUM_GP && S.has_gp &&
         !S.skip_medium && S.mean_type == MUPSI_MEAN_SPHERE && S.gp.ppc <= kMaxPpcCached &&
         (S.n_lights == 0 || S.n_prims == 0);
}

static mupsi_status enqueue_wavefront(mupsi_scene* sc, const PathSrc& src, int64_t n,
                                      const volatile int32_t* abort_flag, EventPairs& evs,
                                      EventPairs& gevs, GuideRun* gr, bool timed = true) {
  mupsi_ctx* ctx = sc->ctx;
  hipStream_t st = ctx->stream;
  PathBuffers& pb = *sc->pb;
  const DevScene& S = sc->ds;
  const int ni = (int)n;
  hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, pb.counters, 8, ni, 0);
  InitArgs ia{pb.P, src, ni, pb.qA[0], nullptr};
  hipLaunchKernelGGL(k_init<false>, dim3(grid_for(n)), dim3(256), 0, st, S, ia);
  HIP_TRY(hipGetLastError());
  mupsi_status status = MUPSI_OK;
  GuideBuffers* gb = gr ? sc->gb.get() : nullptr;
  const int K = gr ? gr->g->cfg.max_records : 0;
  const int iters = std::max(1, S.max_bounce);
  for (int it = 0; it < iters; ++it) {
    if (abort_flag && *abort_flag) {
      status = MUPSI_ERR_ABORTED;
      break;
    }
    const int cur = it & 1, nxt = cur ^ 1;
    int* nA = pb.counters + 2 * cur;
    int* nM = pb.counters + 2 * cur + 1;
    int* nA2 = pb.counters + 2 * nxt;
    int* nM2 = pb.counters + 2 * nxt + 1;
    int* work = pb.counters + 4;
    if (it > 0) {
      // zero the work counter
      hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, work, 1, 0, -1);
      MarchArgs ma{};
      ma.queue = pb.qM[cur];
      ma.count_dev = nM;
      ma.ox = pb.P.ox; ma.oy = pb.P.oy; ma.oz = pb.P.oz;
      ma.dx = pb.P.dx; ma.dy = pb.P.dy; ma.dz = pb.P.dz;
      ma.nt = pb.P.nt;
      ma.ft = nullptr;
      ma.seed = pb.P.mseed;
      ma.flags = pb.P.flags;
      ma.cx = pb.P.cx; ma.cy = pb.P.cy; ma.cz = pb.P.cz; ma.cu = pb.P.cu;
      ma.cgx = pb.P.cgx; ma.cgy = pb.P.cgy; ma.cgz = pb.P.cgz;
      ma.hit = pb.P.mhit;
      ma.t = pb.P.mt; ma.px = pb.P.mpx; ma.py = pb.P.mpy; ma.pz = pb.P.mpz;
      ma.steps = nullptr;
      ma.work = work;
      ma.stats = pb.stats;
      if (fuse_march(sc, gr != nullptr)) {
        ma.fuse = 1;
        ma.rng = pb.P.rng;
        ma.Tx = pb.P.Tx;
        ma.Ty = pb.P.Ty;
        ma.Tz = pb.P.Tz;
        ma.src_W = src.W;
        ma.src_npix = src.npix;
        ma.src_k_begin = src.k_begin;
        ma.src_k_begin_dev = src.k_begin_dev;
        ma.src_xs = src.xs;
        ma.src_ys = src.ys;
        ma.src_ks = src.ks;
      }
      if (timed) {
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        evs.emplace_back(e0, e1);
        HIP_TRY(hipEventRecord(e0, st));
      }
      launch_march(ctx, S, ma, st);
      if (timed) HIP_TRY(hipEventRecord(evs.back().second, st));
      HIP_TRY(hipGetLastError());
      if (gr) {
        hipEvent_t g0, g1;
        HIP_TRY(hipEventCreate(&g0));
        HIP_TRY(hipEventCreate(&g1));
        gevs.emplace_back(g0, g1);
        HIP_TRY(hipEventRecord(g0, st));
        int* nG = pb.counters + 5;
        hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, nG, 1, 0, -1);
        GuidePrepArgs ga{pb.P, pb.qM[cur], nM, gb->qG, nG, gb->feat, gr->record, K, gb->rec_feat,
                         gr->g->box()};
        hipLaunchKernelGGL(k_guide_prep, dim3(grid_for(n)), dim3(256), 0, st, S, ga);
        guide::launch_infer(gr->g, gb->feat, nG, 0, gb->qG, nullptr, gb->mix, st);
        HIP_TRY(hipEventRecord(g1, st));
        HIP_TRY(hipGetLastError());
      }
    }
    // zero next queue counters
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, nA2, 2, 0, -1);
    GuideShade gs{0, 1.0f, nullptr, 0, 1, nullptr};
    if (gr) gs = GuideShade{1, gr->g->cfg.rr_boost, gb->mix, gr->record, K, gb->rec};
    const bool defer = pb.qS != nullptr;
    int* nS = pb.counters + 7;
    if (defer) hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, nS, 1, 0, -1);
    // a mesh mean's hits: their mean and gradient in a kernel of their own
    const bool hit_mean = it > 0 && S.medium_kind == MUPSI_MEDIUM_GP && S.has_gp &&
                          !S.skip_medium && S.mean_type == MUPSI_MEAN_MESH &&
                          S.mean_mesh.n_nodes > 0 && pb.hmean && pb.hbins;
    if (hit_mean) {
      launch_hit_means(st, (int)n, ctx->query_blocks, false, S.mean_mesh, nullptr, nullptr, 1,
                       pb.qM[cur], nM, nullptr, nullptr, pb.P, pb.hbins, pb.capacity, pb.hmean);
      HIP_TRY(hipGetLastError());
    }
    ShadeArgs sa{pb.P, src, pb.qA[cur], nA, pb.qA[nxt], nA2, pb.qM[nxt], nM2, gs,
                 defer ? 1 : 0, pb.SH, pb.qS, nS, nullptr, nullptr, nullptr, nullptr, nullptr,
                 hit_mean ? pb.hmean : nullptr};
    if (scene_has_tri(sc)) hipLaunchKernelGGL((k_shade<false, true>), dim3(grid_for(n)), dim3(256), 0, st, S, sa);
    else hipLaunchKernelGGL((k_shade<false, false>), dim3(grid_for(n)), dim3(256), 0, st, S, sa);
    HIP_TRY(hipGetLastError());
    if (defer) {
      // transmittance of this iteration's medium light samples (gpmedium.cpp:105-129)
      int* work = pb.counters + 4;
      hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, work, 1, 0, -1);
      const ShadowSoA& H = pb.SH;
      MarchArgs ma{};
      ma.queue = pb.qS;
      ma.count_dev = nS;
      ma.ox = H.ox; ma.oy = H.oy; ma.oz = H.oz;
      ma.dx = H.dx; ma.dy = H.dy; ma.dz = H.dz;
      ma.nt = H.nt;
      ma.ft = H.ft;
      ma.seed = H.seed;
      ma.flags = nullptr;
      ma.active = H.active;
      ma.cx = H.cx; ma.cy = H.cy; ma.cz = H.cz; ma.cu = H.cu;
      ma.cgx = H.cgx; ma.cgy = H.cgy; ma.cgz = H.cgz;
      ma.hit = H.hit;
      ma.t = H.t; ma.px = H.px; ma.py = H.py; ma.pz = H.pz;
      ma.steps = nullptr;
      ma.work = work;
      ma.stats = pb.stats;
      if (timed) {
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        evs.emplace_back(e0, e1);
        HIP_TRY(hipEventRecord(e0, st));
      }
      launch_march(ctx, S, ma, st);
      if (timed) HIP_TRY(hipEventRecord(evs.back().second, st));
      NeeArgs na{pb.P, H, pb.qS, nS, gs};
      hipLaunchKernelGGL(k_nee_finish, dim3(grid_for(n)), dim3(256), 0, st, na);
      HIP_TRY(hipGetLastError());
    }
  }
  return status;
}

// Run the wavefront over n paths described by src; leaves per-path radiance in P.L*.
// With sc->use_graph, a plain render (no guide, no abort flag, the pixel grid)
// replays a graph captured from the same sequence of launches.
static mupsi_status run_wavefront(mupsi_scene* sc, const PathSrc& src, int64_t n,
                                  const volatile int32_t* abort_flag, double* march_ms,
                                  int* launches, GuideRun* gr = nullptr) {
  hipStream_t st = sc->ctx->stream;
  if (sc->use_graph && !gr && !abort_flag && !src.xs && !src.plan) {
    if (!sc->graph || sc->graph->n != n) {
      HIP_TRY(hipStreamSynchronize(st));
      sc->graph.reset();
      auto g = std::make_unique<WavefrontGraph>();
      HIP_TRY(g->mem.alloc(1, &g->kb_dev));
      PathSrc gsrc = src;
      gsrc.k_begin_dev = g->kb_dev;
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const mupsi_status cs =
          enqueue_wavefront(sc, gsrc, n, nullptr, g->evs, g->gevs, nullptr, /*timed=*/false);
      hipGraph_t graph = nullptr;
      const hipError_t ce = hipStreamEndCapture(st, &graph);
      if (cs != MUPSI_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return cs;
      }
      HIP_TRY(ce);
      const hipError_t ie = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_TRY(ie);
      g->n = n;
      sc->graph = std::move(g);
      ++sc->graph_builds;
    }
    WavefrontGraph& g = *sc->graph;
    HIP_TRY(hipMemcpyAsync(g.kb_dev, &src.k_begin, sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipGraphLaunch(g.exec, st));
    HIP_TRY(hipStreamSynchronize(st));
    // the replayed launches are not timed one by one (march_ms stays as it is)
    return MUPSI_OK;
  }
  EventPairs evs, gevs;
  mupsi_status status = enqueue_wavefront(sc, src, n, abort_flag, evs, gevs, gr);
  HIP_TRY(hipStreamSynchronize(st));
  if (march_ms) *march_ms += elapsed_ms(evs);
  if (launches) *launches += (int)evs.size();
  if (gr) gr->ms += elapsed_ms(gevs);
  destroy_events(evs);
  destroy_events(gevs);
  return status;
}

extern "C" mupsi_status mupsi_render_passes(mupsi_scene* sc, int32_t k_begin, int32_t k_count,
                                            int32_t passes_per_launch,
                                            const volatile int32_t* abort_flag) {
  if (!sc) return fail(MUPSI_ERR_INVALID, "mupsi_render_passes: null scene");
  if (k_begin < 0 || k_count < 0) return fail(MUPSI_ERR_INVALID, "mupsi_render_passes: negative pass range");
  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(sc->ctx->device));
  const int64_t npix = (int64_t)sc->ds.w * sc->ds.h;
  int ppl = passes_per_launch;
  if (ppl <= 0) ppl = (int)std::max<int64_t>(1, std::min<int64_t>(k_count, (int64_t)(8 << 20) / npix));
  ppl = std::max(1, std::min(ppl, std::max(1, k_count)));
  if ((int64_t)ppl * npix > INT32_MAX) ppl = (int)std::max<int64_t>(1, INT32_MAX / npix);
  mupsi_status st = ensure_paths(sc, (int64_t)ppl * npix);
  if (st != MUPSI_OK) return st;
  mupsi_guide* guide = sc->guide;
  if (guide) {
    const bool will_train = guide->stats.train_passes_done < guide->cfg.train_passes;
    const int64_t per_pass = npix * guide->cfg.max_records * guide->cfg.passes_per_update;
    const int64_t ring = std::min<int64_t>(std::max<int64_t>(per_pass, 8 << 20),
                                           (int64_t)std::max(1, guide->cfg.train_passes) * per_pass);
    st = ensure_guide_buffers(sc, (int64_t)ppl * npix, will_train ? per_pass : 1,
                              will_train ? ring : 1);
    if (st != MUPSI_OK) return st;
  }
  PathBuffers& pb = *sc->pb;
  hipStream_t strm = sc->ctx->stream;
  HIP_TRY(hipMemsetAsync(pb.stats, 0, 8 * sizeof(unsigned long long), strm));
  mupsi_stats stats{};
  double march_ms = 0;
  int launches = 0;
  for (int kb = k_begin; kb < k_begin + k_count;) {
    if (abort_flag && *abort_flag) return fail(MUPSI_ERR_ABORTED, "render aborted");
    // online training: the guide's first train_passes passes are traced
    // passes_per_update at a time, and their recorded vertices train the network
    // before the next ones
    const bool training = guide && guide->stats.train_passes_done < guide->cfg.train_passes;
    const int kc =
        training ? (int)std::min<int64_t>({(int64_t)guide->cfg.passes_per_update,
                                           guide->cfg.train_passes - guide->stats.train_passes_done,
                                           (int64_t)(k_begin + k_count - kb), (int64_t)ppl})
                 : std::min(ppl, k_begin + k_count - kb);
    PathSrc src{sc->ds.w, (int)npix, kb, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
    GuideRun gr{guide, training ? 1 : 0, 0.0};
    st = run_wavefront(sc, src, (int64_t)kc * npix, abort_flag, &march_ms, &launches,
                       guide ? &gr : nullptr);
    if (st == MUPSI_ERR_ABORTED) return fail(st, "render aborted");
    if (st != MUPSI_OK) return st;
    hipLaunchKernelGGL(k_accum, dim3(grid_for(npix)), dim3(256), 0, strm, pb.P.Lx, pb.P.Ly, pb.P.Lz,
                       (int)npix, kc, sc->accum);
    HIP_TRY(hipGetLastError());
    if (guide) guide->stats.infer_ms += gr.ms;
    if (training) {
      GuideBuffers& gb = *sc->gb;
      int* tr_count = pb.counters + 6;
      hipEvent_t t0e, t1e;
      HIP_TRY(hipEventCreate(&t0e));
      HIP_TRY(hipEventCreate(&t1e));
      HIP_TRY(hipEventRecord(t0e, strm));
      hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, strm, tr_count, 1, 0, -1);
      const int64_t np = (int64_t)kc * npix;
      CollectArgs ca{pb.P, (int)np, guide->cfg.max_records, gb.rec, gb.rec_feat, gb.tr_feat,
                     gb.tr_dir, gb.tr_w, tr_count, gb.tr_head, gb.tr_capacity,
                     guide->cfg.indirect_clamp};
      hipLaunchKernelGGL(k_collect, dim3(grid_for(np)), dim3(256), 0, strm, ca);
      HIP_TRY(hipGetLastError());
      int count = 0;
      HIP_TRY(hipMemcpyAsync(&count, tr_count, sizeof(int), hipMemcpyDeviceToHost, strm));
      HIP_TRY(hipStreamSynchronize(strm));
      gb.tr_head = (gb.tr_head + count) % gb.tr_capacity;
      gb.tr_size = std::min(gb.tr_size + count, gb.tr_capacity);
      const int batch = guide->cfg.batch_size;
      // about four passes over this pass's new records, drawn from the whole ring,
      // and at least min_steps_per_pass steps
      const int steps = (int)std::min<int64_t>(
          (int64_t)guide->cfg.max_steps_per_pass * kc,
          std::max<int64_t>((int64_t)guide->cfg.min_steps_per_pass * kc,
                            (4LL * count + batch - 1) / batch));
      if (gb.tr_size > 0 && steps > 0) {
        st = guide::train_random(guide, gb.tr_feat, gb.tr_dir, gb.tr_w, (int)gb.tr_size, steps,
                                 batch, strm);
        if (st != MUPSI_OK) return st;
      }
      HIP_TRY(hipEventRecord(t1e, strm));
      HIP_TRY(hipEventSynchronize(t1e));
      float tms = 0;
      (void)hipEventElapsedTime(&tms, t0e, t1e);
      (void)hipEventDestroy(t0e);
      (void)hipEventDestroy(t1e);
      guide->stats.train_ms += tms;
      guide->stats.train_samples += count;
      guide->stats.train_passes_done += kc;
      if (gb.tr_size > 0 && steps > 0) {
        const int nb = guide::train_blocks(batch);
        std::vector<float> lp((size_t)nb);
        HIP_TRY(hipMemcpy(lp.data(), guide->loss_part, nb * sizeof(float), hipMemcpyDeviceToHost));
        double ls = 0;
        for (int i = 0; i < nb; ++i) ls += lp[i];
        guide->stats.last_loss = ls / batch;
      }
    }
    HIP_TRY(hipStreamSynchronize(strm));
    sc->sample_count += kc;
    stats.paths += (int64_t)kc * npix;
    kb += kc;
  }
  unsigned long long hs[8];
  HIP_TRY(hipMemcpy(hs, pb.stats, sizeof(hs), hipMemcpyDeviceToHost));
  stats.field_evals = (int64_t)hs[0];
  stats.march_steps = (int64_t)hs[1];
  stats.march_calls = (int64_t)hs[2];
  stats.impulse_evals = stats.field_evals * 27 * sc->ds.gp.ppc;
  stats.march_kernel_launches = launches;
  stats.march_kernel_ms = march_ms;
  stats.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  sc->last = stats;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_last_stats(mupsi_scene* sc, mupsi_stats* out) {
  if (!sc || !out) return fail(MUPSI_ERR_INVALID, "mupsi_last_stats: null argument");
  *out = sc->last;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_reset_accum(mupsi_scene* sc) {
  if (!sc) return fail(MUPSI_ERR_INVALID, "mupsi_reset_accum: null scene");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  HIP_TRY(hipMemset(sc->accum, 0, (size_t)sc->ds.w * sc->ds.h * 3 * sizeof(double)));
  sc->sample_count = 0;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_read_accum(mupsi_scene* sc, double* rgb_sum, int64_t* count) {
  if (!sc) return fail(MUPSI_ERR_INVALID, "mupsi_read_accum: null scene");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  if (rgb_sum)
    HIP_TRY(hipMemcpy(rgb_sum, sc->accum, (size_t)sc->ds.w * sc->ds.h * 3 * sizeof(double),
                      hipMemcpyDeviceToHost));
  if (count) *count = sc->sample_count;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_write_accum(mupsi_scene* sc, const double* rgb_sum, int64_t count) {
  if (!sc || !rgb_sum || count < 0) return fail(MUPSI_ERR_INVALID, "mupsi_write_accum: bad argument");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  HIP_TRY(hipMemcpy(sc->accum, rgb_sum, (size_t)sc->ds.w * sc->ds.h * 3 * sizeof(double),
                    hipMemcpyHostToDevice));
  sc->sample_count = count;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_trace_paths(mupsi_scene* sc, const int32_t* x, const int32_t* y,
                                          const int32_t* k, int64_t n, float* rgb_out) {
  if (!sc || (n > 0 && (!x || !y || !k || !rgb_out)))
    return fail(MUPSI_ERR_INVALID, "mupsi_trace_paths: null argument");
  if (n == 0) return MUPSI_OK;
  if (n > INT32_MAX) return fail(MUPSI_ERR_INVALID, "mupsi_trace_paths: too many paths");
  for (int64_t i = 0; i < n; ++i)
    if (x[i] < 0 || y[i] < 0 || k[i] < 0 || x[i] >= sc->ds.w || y[i] >= sc->ds.h)
      return fail(MUPSI_ERR_INVALID, "mupsi_trace_paths: pixel outside the image");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  mupsi_status st = ensure_paths(sc, n);
  if (st != MUPSI_OK) return st;
  DeviceBuffers tmp;
  const int32_t *dx, *dy, *dk;
  HIP_TRY(tmp.upload(x, n, &dx));
  HIP_TRY(tmp.upload(y, n, &dy));
  HIP_TRY(tmp.upload(k, n, &dk));
  PathBuffers& pb = *sc->pb;
  HIP_TRY(hipMemsetAsync(pb.stats, 0, 8 * sizeof(unsigned long long), sc->ctx->stream));
  PathSrc src{sc->ds.w, sc->ds.w * sc->ds.h, 0, nullptr, dx, dy, dk, nullptr, 0, nullptr};
  st = run_wavefront(sc, src, n, nullptr, nullptr, nullptr);
  if (st != MUPSI_OK) return st;
  float* drgb;
  HIP_TRY(tmp.alloc(3 * n, &drgb));
  hipLaunchKernelGGL(k_interleave3, dim3(grid_for(n)), dim3(256), 0, sc->ctx->stream, pb.P.Lx,
                     pb.P.Ly, pb.P.Lz, n, drgb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(sc->ctx->stream));
  HIP_TRY(hipMemcpy(rgb_out, drgb, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
  return MUPSI_OK;
}

// ====================================================================== host: batched scenes

struct mupsi_batch {
  mupsi_ctx* ctx = nullptr;
  std::vector<mupsi_scene*> scenes;
  bool valid = true;  // false once a member scene was freed
  DeviceBuffers mem;
  DevScene* d_scenes = nullptr;
  CachedConsts* d_cc = nullptr;
  LaunchPlan* d_plan = nullptr;
  int64_t* d_total = nullptr;
  std::vector<LaunchPlan> plan;
  std::unique_ptr<PathBuffers> pb;
  int max_bounce = 1;
  bool any_cached = false, any_plain = false;
  bool any_cached_mesh = false;  // window-marcher scenes with a mean mesh
  bool any_mesh_mean = false;    // a GP medium with a mean mesh (k_hit_mean serves its hits)
  bool any_tri = false;          // a scene holds a triangle mesh (primitive or mean)
  mupsi_batch_tuning tuning{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t graph_cap = 0;
  int graph_blocks = 0;
  mupsi_batch_stats stats{};
  ~mupsi_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// nA0 = the launch's path count (all paths start in queue A); every other counter 0.
__global__ void k_batch_counters(int* c, const int64_t* total) {
  const int i = threadIdx.x;
  if (i < 16) c[i] = i == 0 ? (int)*total : 0;
}

static bool scene_uses_cached(const DevScene& S) {
  return S.medium_kind == MUPSI_MEDIUM_GP && S.gp.ppc <= kMaxPpcCached;
}

// The kernels of one batched launch (paths of the device plan), on stream st.
// Eager or under graph capture: nothing here synchronises or touches the host.
static void enqueue_batch_launch(mupsi_batch* b, hipStream_t st) {
  PathBuffers& pb = *b->pb;
  const int64_t n = pb.capacity;
  mupsi_ctx* ctx = b->ctx;
  PathSrc src{0, 1, 0, nullptr, nullptr, nullptr, nullptr, b->d_plan, (int)b->scenes.size(), b->d_total};
  hipLaunchKernelGGL(k_batch_counters, dim3(1), dim3(64), 0, st, pb.counters, b->d_total);
  InitArgs ia{pb.P, src, (int)n, pb.qA[0], b->d_scenes};
  hipLaunchKernelGGL(k_init<true>, dim3(grid_for(n)), dim3(256), 0, st, DevScene{}, ia);
  const int per_cu = std::max(1, b->tuning.march_blocks_per_cu);
  for (int it = 0; it < b->max_bounce; ++it) {
    const int cur = it & 1, nxt = cur ^ 1;
    int* nA = pb.counters + 2 * cur;
    int* nM = pb.counters + 2 * cur + 1;
    int* nMm = pb.counters + 8 + cur;
    int* work = pb.counters + 4;
    int* work2 = pb.counters + 10;
    int* nMs = pb.counters + 12 + cur;
    int* work3 = pb.counters + 14;
    if (it > 0) {
      MarchArgs ma{};
      ma.ox = pb.P.ox; ma.oy = pb.P.oy; ma.oz = pb.P.oz;
      ma.dx = pb.P.dx; ma.dy = pb.P.dy; ma.dz = pb.P.dz;
      ma.nt = pb.P.nt;
      ma.ft = nullptr;
      ma.seed = pb.P.mseed;
      ma.flags = pb.P.flags;
      ma.cx = pb.P.cx; ma.cy = pb.P.cy; ma.cz = pb.P.cz; ma.cu = pb.P.cu;
      ma.cgx = pb.P.cgx; ma.cgy = pb.P.cgy; ma.cgz = pb.P.cgz;
      ma.hit = pb.P.mhit;
      ma.t = pb.P.mt; ma.px = pb.P.mpx; ma.py = pb.P.mpy; ma.pz = pb.P.mpz;
      ma.stats = pb.stats;
      ma.scenes = b->d_scenes;
      ma.sid = pb.P.sid;
      const dim3 wgrid(per_cu * ctx->cus * 2), wblock(64);
      if (b->any_cached) {
        hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, work, 1, 0, -1);
        ma.queue = pb.qM[cur];
        ma.count_dev = nM;
        ma.work = work;
        hipLaunchKernelGGL((k_march_window<kMaxPpcCached, false, true, kMultiGroup>), wgrid,
                           wblock, 0, st, ma, DevScene{}, CachedConsts{}, b->d_cc);
      }
      if (b->any_cached_mesh) {
        hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, work3, 1, 0, -1);
        ma.queue = pb.qMs[cur];
        ma.count_dev = nMs;
        ma.work = work3;
        hipLaunchKernelGGL((k_march_window<kMaxPpcCached, true, true, kMultiGroup>), wgrid,
                           wblock, 0, st, ma, DevScene{}, CachedConsts{}, b->d_cc);
      }
      if (b->any_plain) {
        hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, work2, 1, 0, -1);
        ma.queue = pb.qMm[cur];
        ma.count_dev = nMm;
        ma.work = work2;
        hipLaunchKernelGGL(k_march<true>, dim3(per_cu * ctx->cus), dim3(256), 0, st, DevScene{}, ma);
      }
    }
    int* nA2 = pb.counters + 2 * nxt;
    int* nMm2 = pb.counters + 8 + nxt;
    int* nMs2 = pb.counters + 12 + nxt;
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, nA2, 2, 0, -1);
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, nMm2, 1, 0, -1);
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, nMs2, 1, 0, -1);
    GuideShade gs{0, 1.0f, nullptr, 0, 1, nullptr};
    // the mesh-mean hits of this iteration's marches (window marcher of mesh means,
    // lane-per-ray marcher): their mean and gradient in a kernel of their own
    const bool hit_mean = it > 0 && b->any_mesh_mean && pb.hmean && pb.hbins &&
                          (b->any_cached_mesh || b->any_plain);
    if (hit_mean) {
      // both hit queues in one sort: window-marcher (Ms) then lane-per-ray (Mm) paths
      const int* qa = b->any_cached_mesh ? pb.qMs[cur] : pb.qMm[cur];
      const int* na = b->any_cached_mesh ? pb.counters + 12 + cur : pb.counters + 8 + cur;
      const bool both = b->any_cached_mesh && b->any_plain;
      launch_hit_means(st, (int)n, ctx->query_blocks, true, DevMesh{}, b->d_scenes, pb.P.sid,
                       (int)b->scenes.size(), qa, na, both ? pb.qMm[cur] : nullptr,
                       both ? pb.counters + 8 + cur : nullptr, pb.P, pb.hbins, pb.capacity,
                       pb.hmean);
    }
    ShadeArgs sa{pb.P, src, pb.qA[cur], nA, pb.qA[nxt], nA2, pb.qM[nxt], pb.counters + 2 * nxt + 1,
                 gs, 0, ShadowSoA{}, nullptr, nullptr, b->d_scenes,
                 pb.qMm[nxt], nMm2, pb.qMs[nxt], nMs2, hit_mean ? pb.hmean : nullptr};
    if (b->any_tri) hipLaunchKernelGGL((k_shade<true, true>), dim3(grid_for(n)), dim3(256), 0, st, DevScene{}, sa);
    else hipLaunchKernelGGL((k_shade<true, false>), dim3(grid_for(n)), dim3(256), 0, st, DevScene{}, sa);
  }
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    const mupsi_scene* sc = b->scenes[i];
    const int npix = sc->ds.w * sc->ds.h;
    hipLaunchKernelGGL(k_accum_plan, dim3(grid_for(npix)), dim3(256), 0, st, pb.P.Lx, pb.P.Ly,
                       pb.P.Lz, b->d_plan, (int)i, sc->accum);
  }
}

static mupsi_status batch_buffers(mupsi_batch* b, int64_t cap) {
  if (b->pb && b->pb->capacity == cap) return MUPSI_OK;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return alloc_path_buffers(b->pb, cap, false, b->any_plain || b->any_cached_mesh, b->any_mesh_mean);
}

// Capture the launch's kernels into a graph for the current buffers and tuning.
static mupsi_status batch_graph(mupsi_batch* b) {
  const int blocks = std::max(1, b->tuning.march_blocks_per_cu);
  if (b->exec && b->graph_cap == b->pb->capacity && b->graph_blocks == blocks) return MUPSI_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  hipStream_t st = b->ctx->stream;
  HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  enqueue_batch_launch(b, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (e != hipSuccess) return fail(MUPSI_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_cap = b->pb->capacity;
  b->graph_blocks = blocks;
  b->stats.graph_builds += 1;
  return MUPSI_OK;
}

// Fill the next launch: every unfinished scene gets passes up to a fair share of
// `cap` paths (at least one pass while it fits), in a rotating order; offsets
// are assigned in scene order, as the plan search in k_init expects.
static int64_t batch_plan(mupsi_batch* b, std::vector<int>& next_k, std::vector<int>& left,
                          int64_t cap, size_t rotate) {
  const size_t ns = b->scenes.size();
  std::vector<int> kc(ns, 0);
  size_t active = 0;
  for (size_t s = 0; s < ns; ++s) active += left[s] > 0;
  if (active == 0) return 0;
  const int64_t fair = std::max<int64_t>(1, cap / (int64_t)active);
  int64_t used = 0;
  for (size_t i = 0; i < ns; ++i) {
    const size_t s = (rotate + i) % ns;
    if (left[s] <= 0) continue;
    const int64_t npix = (int64_t)b->scenes[s]->ds.w * b->scenes[s]->ds.h;
    int64_t want = std::min<int64_t>(left[s], std::max<int64_t>(1, fair / npix));
    want = std::min<int64_t>(want, (cap - used) / npix);
    if (want <= 0) continue;
    kc[s] = (int)want;
    used += want * npix;
  }
  int64_t off = 0;
  for (size_t s = 0; s < ns; ++s) {
    const mupsi_scene* sc = b->scenes[s];
    const int npix = sc->ds.w * sc->ds.h;
    b->plan[s] = LaunchPlan{off, npix, sc->ds.w, next_k[s], kc[s]};
    off += (int64_t)kc[s] * npix;
    next_k[s] += kc[s];
    left[s] -= kc[s];
  }
  return off;
}

static mupsi_status batch_launch(mupsi_batch* b, int64_t total, double* ms) {
  hipStream_t st = b->ctx->stream;
  HIP_TRY(hipMemcpyAsync(b->d_plan, b->plan.data(), b->plan.size() * sizeof(LaunchPlan),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(b->d_total, &total, sizeof(int64_t), hipMemcpyHostToDevice, st));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, st));
  if (b->tuning.use_graph) {
    mupsi_status s2 = batch_graph(b);
    if (s2 != MUPSI_OK) return s2;
    HIP_TRY(hipGraphLaunch(b->exec, st));
  } else {
    enqueue_batch_launch(b, st);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, st));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  for (size_t s = 0; s < b->scenes.size(); ++s) b->scenes[s]->sample_count += b->plan[s].kc;
  b->stats.paths += total;
  b->stats.launches += 1;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_batch_create(mupsi_ctx* ctx, mupsi_scene* const* scenes, int32_t n,
                                           mupsi_batch** out) {
  if (!ctx || !scenes || !out || n <= 0) return fail(MUPSI_ERR_INVALID, "mupsi_batch_create: bad argument");
  if (n > 65535) return fail(MUPSI_ERR_INVALID, "mupsi_batch_create: at most 65535 scenes");
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  auto b = std::make_unique<mupsi_batch>();
  b->ctx = ctx;
  std::vector<DevScene> ds;
  std::vector<CachedConsts> cc;
  for (int i = 0; i < n; ++i) {
    mupsi_scene* sc = scenes[i];
    if (!sc || sc->ctx != ctx) return fail(MUPSI_ERR_INVALID, "mupsi_batch_create: scene of another context");
    b->scenes.push_back(sc);
    ds.push_back(sc->ds);
    cc.push_back(cached_consts(sc->ds.gp));
    b->max_bounce = std::max(b->max_bounce, sc->ds.max_bounce);
    if (scene_has_tri(sc)) b->any_tri = true;
    if (sc->ds.has_gp && !sc->ds.skip_medium) {
      if (!scene_uses_cached(sc->ds)) b->any_plain = true;
      else if (sc->ds.mean_type == MUPSI_MEAN_MESH) b->any_cached_mesh = true;
      else b->any_cached = true;
      if (sc->ds.medium_kind == MUPSI_MEDIUM_GP && sc->ds.mean_type == MUPSI_MEAN_MESH)
        b->any_mesh_mean = true;
    }
  }
  const DevScene* dds;
  HIP_TRY(b->mem.upload(ds.data(), ds.size(), &dds));
  b->d_scenes = const_cast<DevScene*>(dds);
  const CachedConsts* dcc;
  HIP_TRY(b->mem.upload(cc.data(), cc.size(), &dcc));
  b->d_cc = const_cast<CachedConsts*>(dcc);
  HIP_TRY(b->mem.alloc(n, &b->d_plan));
  HIP_TRY(b->mem.alloc(1, &b->d_total));
  b->plan.assign(n, LaunchPlan{});
  b->tuning.use_graph = 1;
  ctx->batches.push_back(b.get());
  *out = b.release();
  return MUPSI_OK;
}

static void destroy_batch(mupsi_batch* b) {
  (void)hipStreamSynchronize(b->ctx->stream);
  delete b;
}

// A freed scene invalidates every batch that holds it.
static void batch_forget_scene(mupsi_ctx* ctx, const mupsi_scene* scene) {
  for (mupsi_batch* b : ctx->batches)
    if (std::find(b->scenes.begin(), b->scenes.end(), scene) != b->scenes.end()) b->valid = false;
}

static bool scene_in_live_batch(const mupsi_ctx* ctx, const mupsi_scene* scene) {
  for (const mupsi_batch* b : ctx->batches)
    if (b->valid && std::find(b->scenes.begin(), b->scenes.end(), scene) != b->scenes.end())
      return true;
  return false;
}

extern "C" void mupsi_batch_free(mupsi_batch* b) {
  if (!b) return;
  mupsi_ctx* ctx = b->ctx;
  auto it = std::find(ctx->batches.begin(), ctx->batches.end(), b);
  if (it != ctx->batches.end()) ctx->batches.erase(it);
  (void)hipSetDevice(ctx->device);
  destroy_batch(b);
}

extern "C" mupsi_status mupsi_batch_set_tuning(mupsi_batch* b, const mupsi_batch_tuning* t) {
  if (!b || !t) return fail(MUPSI_ERR_INVALID, "mupsi_batch_set_tuning: null argument");
  if (t->paths_per_launch < 0 || t->march_blocks_per_cu < 0 || t->march_blocks_per_cu > 64)
    return fail(MUPSI_ERR_INVALID, "mupsi_batch_set_tuning: bad tuning");
  b->tuning = *t;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_batch_get_tuning(mupsi_batch* b, mupsi_batch_tuning* t) {
  if (!b || !t) return fail(MUPSI_ERR_INVALID, "mupsi_batch_get_tuning: null argument");
  *t = b->tuning;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_batch_stats_get(mupsi_batch* b, mupsi_batch_stats* out) {
  if (!b || !out) return fail(MUPSI_ERR_INVALID, "mupsi_batch_stats_get: null argument");
  *out = b->stats;
  return MUPSI_OK;
}

// Autotuning.  Larger launches amortise the tail of each bounce loop (few live
// paths, idle CUs): measured on the generated 64-scene workload, 16M paths per
// launch beat 8M, 4M and 2M at every marcher grid.  The launch size is therefore
// the whole remaining work up to kTuneCap (about 2.5 GB of path state); only the
// persistent marchers' blocks per CU are probed, on real passes, when the work
// spans enough launches to pay for the probes.
static constexpr int64_t kTuneCap = 1LL << 24;
static const int kTuneBlocks[] = {8, 16, 4};

extern "C" mupsi_status mupsi_batch_render(mupsi_batch* b, const int32_t* k_begin,
                                           const int32_t* k_count,
                                           const volatile int32_t* abort_flag) {
  if (!b || !k_begin || !k_count) return fail(MUPSI_ERR_INVALID, "mupsi_batch_render: null argument");
  if (!b->valid) return fail(MUPSI_ERR_INVALID, "mupsi_batch_render: a scene of the batch was freed");
  const size_t ns = b->scenes.size();
  std::vector<int> next_k(ns), left(ns);
  int64_t work = 0;
  int64_t min_fit = 0;  // the largest single pass: a launch must hold it
  for (size_t s = 0; s < ns; ++s) {
    if (k_begin[s] < 0 || k_count[s] < 0) return fail(MUPSI_ERR_INVALID, "mupsi_batch_render: negative pass range");
    next_k[s] = k_begin[s];
    left[s] = k_count[s];
    const int64_t npix = (int64_t)b->scenes[s]->ds.w * b->scenes[s]->ds.h;
    work += npix * k_count[s];
    if (k_count[s] > 0) min_fit = std::max(min_fit, npix);
  }
  if (work == 0) return MUPSI_OK;
  HIP_TRY(hipSetDevice(b->ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  size_t rotate = 0;
  // autotuning: launch size from the work, blocks per CU probed on real passes
  if (b->tuning.paths_per_launch == 0 || b->tuning.march_blocks_per_cu == 0) {
    const auto tt0 = std::chrono::steady_clock::now();
    if (b->tuning.paths_per_launch == 0)
      b->tuning.paths_per_launch = std::max(min_fit, std::min(work, kTuneCap));
    if (b->tuning.march_blocks_per_cu == 0) {
      const int64_t cap = b->tuning.paths_per_launch;
      const int nb = (int)(sizeof(kTuneBlocks) / sizeof(kTuneBlocks[0]));
      b->tuning.march_blocks_per_cu = kTuneBlocks[0];
      if (work >= (int64_t)(2 * nb + 2) * cap) {
        mupsi_status st = batch_buffers(b, cap);
        if (st != MUPSI_OK) return st;
        double best = 1e300;
        int best_blk = kTuneBlocks[0];
        for (int blk : kTuneBlocks) {
          if (abort_flag && *abort_flag) return fail(MUPSI_ERR_ABORTED, "render aborted");
          b->tuning.march_blocks_per_cu = blk;
          double per = 0.0;
          // a first launch on a new graph includes its capture: time a second one
          for (int rep = 0; rep < 2; ++rep) {
            const int64_t total = batch_plan(b, next_k, left, cap, rotate++);
            if (total == 0) break;
            double ms = 0;
            st = batch_launch(b, total, &ms);
            if (st != MUPSI_OK) return st;
            per = ms / (double)total;
          }
          if (per > 0.0 && per < best) {
            best = per;
            best_blk = blk;
          }
        }
        b->tuning.march_blocks_per_cu = best_blk;
      }
    }
    b->stats.tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tt0).count();
  }
  const int64_t cap = std::max(b->tuning.paths_per_launch, min_fit);
  mupsi_status st = batch_buffers(b, cap);
  if (st != MUPSI_OK) return st;
  while (true) {
    if (abort_flag && *abort_flag) return fail(MUPSI_ERR_ABORTED, "render aborted");
    const int64_t total = batch_plan(b, next_k, left, cap, rotate++);
    if (total == 0) break;
    st = batch_launch(b, total, nullptr);
    if (st != MUPSI_OK) return st;
  }
  b->stats.total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return MUPSI_OK;
}

// ====================================================================== host: batch API

static DevGP gp_from_params(const mupsi_gp_params& p) { return make_dev_gp(p); }

extern "C" mupsi_status mupsi_guide_sample(mupsi_guide* g, const float* mixture, const float* nrm,
                                           const float* u, int64_t n, float* dir, float* pdf,
                                           float* weight) {
  if (!g) return fail(MUPSI_ERR_INVALID, "mupsi_guide_sample: null guide");
  if (n <= 0) return MUPSI_OK;
  if (!mixture || !nrm || !u || !dir || !pdf || !weight)
    return fail(MUPSI_ERR_INVALID, "mupsi_guide_sample: null array");
  HIP_TRY(hipSetDevice(g->ctx->device));
  hipLaunchKernelGGL(k_guide_sample, dim3(grid_for(n)), dim3(256), 0, g->ctx->stream, mixture, nrm,
                     u, n, dir, pdf, weight);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_psi_eval(mupsi_ctx* ctx, const mupsi_gp_params* gp, const float* px,
                                       const float* py, const float* pz, const uint32_t* seed,
                                       int64_t n, float* out4, float* l1) {
  if (!ctx || !gp) return fail(MUPSI_ERR_INVALID, "mupsi_psi_eval: null argument");
  std::string err;
  if (!validate_gp(*gp, &err)) return fail(MUPSI_ERR_INVALID, err);
  if (n <= 0) return MUPSI_OK;
  if (!px || !py || !pz || !seed || !out4) return fail(MUPSI_ERR_INVALID, "mupsi_psi_eval: null array");
  HIP_TRY(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_psi, dim3(grid_for(n)), dim3(256), 0, ctx->stream, gp_from_params(*gp), px,
                     py, pz, seed, n, out4, l1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_cell_impulses(mupsi_ctx* ctx, const mupsi_gp_params* gp,
                                            const int32_t* cell3, const uint32_t* seed, int64_t n,
                                            uint32_t* cell_seed, float* pos, int8_t* w) {
  if (!ctx || !gp) return fail(MUPSI_ERR_INVALID, "mupsi_cell_impulses: null argument");
  std::string err;
  if (!validate_gp(*gp, &err)) return fail(MUPSI_ERR_INVALID, err);
  if (n <= 0) return MUPSI_OK;
  if (!cell3 || !seed || !cell_seed || !pos || !w)
    return fail(MUPSI_ERR_INVALID, "mupsi_cell_impulses: null array");
  HIP_TRY(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_cell_impulses, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                     gp_from_params(*gp), cell3, seed, n, cell_seed, pos, w);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MUPSI_OK;
}

static CondSoA cond_soa(const mupsi_cond_soa* c) {
  CondSoA r{};
  if (c) r = CondSoA{c->active, c->cx, c->cy, c->cz, c->u_tilde, c->gx, c->gy, c->gz};
  return r;
}
static bool medium_ready(mupsi_scene* sc) { return sc && sc->ds.has_gp; }
static bool gp_ready(mupsi_scene* sc) {
  return medium_ready(sc) && sc->ds.medium_kind == MUPSI_MEDIUM_GP;
}

extern "C" mupsi_status mupsi_field_eval(mupsi_scene* sc, const float* px, const float* py,
                                         const float* pz, const uint32_t* seed,
                                         const mupsi_cond_soa* cond, int64_t n, float* out4) {
  if (!medium_ready(sc)) return fail(MUPSI_ERR_INVALID, "mupsi_field_eval: scene has no medium");
  if (n <= 0) return MUPSI_OK;
  if (!px || !py || !pz || !seed || !out4) return fail(MUPSI_ERR_INVALID, "mupsi_field_eval: null array");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  hipLaunchKernelGGL(k_field, dim3(grid_for(n)), dim3(256), 0, sc->ctx->stream, sc->ds, px, py, pz,
                     seed, cond_soa(cond), n, out4);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(sc->ctx->stream));
  return MUPSI_OK;
}

// Device buffers of one batch march beyond its inputs and outputs.
struct MarchScratch {
  int* work;                     // persistent marcher's ray counter
  unsigned long long* stats;     // 4 counters
  float *ox, *oy, *oz, *dx, *dy, *dz;  // normalized copies of the rays
  float *px, *py, *pz;           // hit points before interleaving
};

static mupsi_status march_check(mupsi_scene* sc, int32_t mode, const float* ox, const float* oy,
                                const float* oz, const float* dx, const float* dy,
                                const float* dz, const float* near_t, const float* far_t,
                                const uint32_t* seed, int64_t n, uint8_t* hit, float* t,
                                float* p, const char* who) {
  if (!medium_ready(sc)) return fail(MUPSI_ERR_INVALID, std::string(who) + ": scene has no medium");
  if (mode != 0 && mode != 1) return fail(MUPSI_ERR_INVALID, std::string(who) + ": mode must be 0 or 1");
  if (n <= 0) return MUPSI_OK;
  if (n > INT32_MAX) return fail(MUPSI_ERR_INVALID, std::string(who) + ": too many rays");
  if (!ox || !oy || !oz || !dx || !dy || !dz || !near_t || !far_t || !seed || !hit || !t || !p)
    return fail(MUPSI_ERR_INVALID, std::string(who) + ": null array");
  return MUPSI_OK;
}

// The march of mupsi_march / mupsi_march_noalloc on the context stream with the
// given scratch; allocates nothing.  The counters are cleared by a fill kernel.
static mupsi_status march_run(mupsi_scene* sc, int32_t mode, const float* ox, const float* oy,
                              const float* oz, const float* dx, const float* dy, const float* dz,
                              const float* near_t, const float* far_t, const uint32_t* seed,
                              const mupsi_cond_soa* cond, int64_t n, uint8_t* hit, float* t,
                              float* p, float* nrm, int32_t* steps, const MarchScratch& ws) {
  hipStream_t st = sc->ctx->stream;
  // work (1 int) and stats (4 x u64 from byte 8): 10 words from the work counter
  static_assert(sizeof(unsigned long long) == 8, "stats words");
  hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, ws.work, 1, 0, -1);
  hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, (int*)ws.stats, 8, 0, -1);
  // Ray(o, d) normalizes the direction; the batch copies o and normalized d.
  hipLaunchKernelGGL(k_normalize_rays, dim3(grid_for(n)), dim3(256), 0, st, ox, oy, oz, dx, dy, dz,
                     n, ws.ox, ws.oy, ws.oz, ws.dx, ws.dy, ws.dz);
  CondSoA cs = cond_soa(cond);
  MarchArgs ma{};
  ma.queue = nullptr;
  ma.count_dev = nullptr;
  ma.count_host = (int)n;
  ma.ox = ws.ox; ma.oy = ws.oy; ma.oz = ws.oz; ma.dx = ws.dx; ma.dy = ws.dy; ma.dz = ws.dz;
  ma.nt = near_t;
  ma.ft = far_t;
  ma.seed = seed;
  ma.flags = nullptr;
  ma.active = cs.active;
  ma.cx = cs.cx; ma.cy = cs.cy; ma.cz = cs.cz; ma.cu = cs.u; ma.cgx = cs.gx; ma.cgy = cs.gy; ma.cgz = cs.gz;
  ma.hit = hit;
  ma.t = t;
  ma.px = ws.px; ma.py = ws.py; ma.pz = ws.pz;
  ma.steps = steps;
  ma.work = ws.work;
  ma.stats = ws.stats;
  launch_march(sc->ctx, sc->ds, ma, st);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_interleave3, dim3(grid_for(n)), dim3(256), 0, st, ws.px, ws.py, ws.pz, n, p);
  if (mode == 0 && nrm)
    hipLaunchKernelGGL(k_hit_normals, dim3(grid_for(n)), dim3(256), 0, st, sc->ds, hit, ws.px,
                       ws.py, ws.pz, seed, cs, n, nrm);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_march(mupsi_scene* sc, int32_t mode, const float* ox,
                                    const float* oy, const float* oz, const float* dx,
                                    const float* dy, const float* dz, const float* near_t,
                                    const float* far_t, const uint32_t* seed,
                                    const mupsi_cond_soa* cond, int64_t n, uint8_t* hit, float* t,
                                    float* p, float* nrm, int32_t* steps) {
  const mupsi_status chk = march_check(sc, mode, ox, oy, oz, dx, dy, dz, near_t, far_t, seed, n,
                                       hit, t, p, "mupsi_march");
  if (chk != MUPSI_OK || n <= 0) return chk;
  HIP_TRY(hipSetDevice(sc->ctx->device));
  DeviceBuffers tmp;
  MarchScratch ws{};
  HIP_TRY(tmp.alloc(1, &ws.work));
  HIP_TRY(tmp.alloc(4, &ws.stats));
  for (float** q : {&ws.ox, &ws.oy, &ws.oz, &ws.dx, &ws.dy, &ws.dz, &ws.px, &ws.py, &ws.pz})
    HIP_TRY(tmp.alloc(n, q));
  return march_run(sc, mode, ox, oy, oz, dx, dy, dz, near_t, far_t, seed, cond, n, hit, t, p, nrm,
                   steps, ws);
}

extern "C" mupsi_status mupsi_march_noalloc(mupsi_scene* sc, int32_t mode, const float* ox,
                                            const float* oy, const float* oz, const float* dx,
                                            const float* dy, const float* dz,
                                            const float* near_t, const float* far_t,
                                            const uint32_t* seed, const mupsi_cond_soa* cond,
                                            int64_t n, uint8_t* hit, float* t, float* p,
                                            float* nrm, int32_t* steps) {
  const mupsi_status chk = march_check(sc, mode, ox, oy, oz, dx, dy, dz, near_t, far_t, seed, n,
                                       hit, t, p, "mupsi_march_noalloc");
  if (chk != MUPSI_OK || n <= 0) return chk;
  if (n > kNoAllocMaxRays)
    return fail(MUPSI_ERR_INVALID, "mupsi_march_noalloc: " + std::to_string(n) +
                                       " rays exceed the scratch maximum of " +
                                       std::to_string(kNoAllocMaxRays));
  HIP_TRY(hipSetDevice(sc->ctx->device));
  char* base = static_cast<char*>(sc->ctx->scratch);
  if (!base) return fail(MUPSI_ERR_HIP, "mupsi_march_noalloc: no scratch address for the device");
  float* arr = reinterpret_cast<float*>(base + kScratchHead);
  MarchScratch ws{};
  ws.work = reinterpret_cast<int*>(base);
  ws.stats = reinterpret_cast<unsigned long long*>(base + 8);
  float** slots[9] = {&ws.ox, &ws.oy, &ws.oz, &ws.dx, &ws.dy, &ws.dz, &ws.px, &ws.py, &ws.pz};
  for (int i = 0; i < 9; ++i) *slots[i] = arr + (size_t)i * kNoAllocMaxRays;
  return march_run(sc, mode, ox, oy, oz, dx, dy, dz, near_t, far_t, seed, cond, n, hit, t, p, nrm,
                   steps, ws);
}

extern "C" mupsi_status mupsi_condition(mupsi_scene* sc, const float* cx, const float* cy,
                                        const float* cz, const float* nx, const float* ny,
                                        const float* nz, const uint32_t* seed_next, int64_t n,
                                        double* u, double* g) {
  if (!gp_ready(sc)) return fail(MUPSI_ERR_INVALID, "mupsi_condition: scene has no GP medium");
  if (n <= 0) return MUPSI_OK;
  if (!cx || !cy || !cz || !nx || !ny || !nz || !seed_next || !u || !g)
    return fail(MUPSI_ERR_INVALID, "mupsi_condition: null array");
  HIP_TRY(hipSetDevice(sc->ctx->device));
  hipLaunchKernelGGL(k_condition, dim3(grid_for(n)), dim3(256), 0, sc->ctx->stream, sc->ds, cx, cy,
                     cz, nx, ny, nz, seed_next, n, u, g);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(sc->ctx->stream));
  return MUPSI_OK;
}

static const DevMesh* pick_mesh(mupsi_scene* sc, int32_t prim) {
  if (!sc) return nullptr;
  if (prim == -1) return sc->host.has_mean_mesh ? &sc->mean_mesh : nullptr;
  if (prim < 0 || prim >= (int)sc->host.prims.size()) return nullptr;
  const DevPrim& pr = sc->host.prims[prim];
  if (pr.type != MUPSI_PRIM_MESH) return nullptr;
  return &sc->meshes[pr.mesh];
}

// Grid of the mesh-query kernels that copy the top of the tree into LDS: one
// thread per query up to blocks that are all resident at once (a block beyond
// them would start only when a resident one ends, after its share of the
// queries), each block looping over its queries.
static unsigned query_grid(const mupsi_ctx* ctx, int64_t n) {
  return std::min<unsigned>(grid_for(n), (unsigned)std::max(1, ctx->query_blocks));
}

// Launches one mesh-query kernel on the context's stream, waits for it, and
// keeps its device time (HIP events on that stream) for mupsi_last_kernel_ms.
template <class F>
static mupsi_status mesh_query_launch(mupsi_scene* sc, F&& launch) {
  mupsi_ctx* ctx = sc->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipEventRecord(ctx->ev_begin, ctx->stream));
  launch(ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ctx->ev_end, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipEventElapsedTime(&ctx->last_kernel_ms, ctx->ev_begin, ctx->ev_end));
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_last_kernel_ms(mupsi_ctx* ctx, float* ms) {
  if (!ctx || !ms) return fail(MUPSI_ERR_INVALID, "mupsi_last_kernel_ms: null argument");
  *ms = ctx->last_kernel_ms;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_mesh_intersect(mupsi_scene* sc, int32_t prim, const float* ox,
                                             const float* oy, const float* oz, const float* dx,
                                             const float* dy, const float* dz,
                                             const float* near_t, const float* far_t, int64_t n,
                                             int32_t* tri, float* t, float* u, float* v) {
  const DevMesh* m = pick_mesh(sc, prim);
  if (!m) return fail(MUPSI_ERR_INVALID, "mupsi_mesh_intersect: no such mesh");
  if (n <= 0) return MUPSI_OK;
  if (!ox || !oy || !oz || !dx || !dy || !dz || !near_t || !far_t || !tri || !t || !u || !v)
    return fail(MUPSI_ERR_INVALID, "mupsi_mesh_intersect: null buffer");
  RaySoA rs{ox, oy, oz, dx, dy, dz, near_t, far_t};
  return mesh_query_launch(sc, [&](hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_intersect, dim3(query_grid(sc->ctx, n)), dim3(256), 0, st, *m, rs, n, tri, t,
                       u, v);
  });
}

extern "C" mupsi_status mupsi_mesh_occluded(mupsi_scene* sc, int32_t prim, const float* ox,
                                            const float* oy, const float* oz, const float* dx,
                                            const float* dy, const float* dz, const float* near_t,
                                            const float* far_t, int64_t n, uint8_t* occ) {
  const DevMesh* m = pick_mesh(sc, prim);
  if (!m) return fail(MUPSI_ERR_INVALID, "mupsi_mesh_occluded: no such mesh");
  if (n <= 0) return MUPSI_OK;
  if (!ox || !oy || !oz || !dx || !dy || !dz || !near_t || !far_t || !occ)
    return fail(MUPSI_ERR_INVALID, "mupsi_mesh_occluded: null buffer");
  RaySoA rs{ox, oy, oz, dx, dy, dz, near_t, far_t};
  return mesh_query_launch(sc, [&](hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_occluded, dim3(query_grid(sc->ctx, n)), dim3(256), 0, st, *m, rs, n, occ);
  });
}

extern "C" mupsi_status mupsi_mesh_closest(mupsi_scene* sc, int32_t prim, const float* px,
                                           const float* py, const float* pz, int64_t n,
                                           float* dist, int32_t* tri, float* closest3) {
  const DevMesh* m = pick_mesh(sc, prim);
  if (!m) return fail(MUPSI_ERR_INVALID, "mupsi_mesh_closest: no such mesh");
  if (n <= 0) return MUPSI_OK;
  if (!px || !py || !pz || !dist || !tri || !closest3)
    return fail(MUPSI_ERR_INVALID, "mupsi_mesh_closest: null buffer");
  return mesh_query_launch(sc, [&](hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_closest, dim3(query_grid(sc->ctx, n)), dim3(256), 0, st, *m, px, py, pz, n,
                       dist, tri, closest3);
  });
}

extern "C" mupsi_status mupsi_mesh_inside(mupsi_scene* sc, int32_t prim, const float* px,
                                          const float* py, const float* pz, int64_t n,
                                          uint8_t* inside) {
  const DevMesh* m = pick_mesh(sc, prim);
  if (!m) return fail(MUPSI_ERR_INVALID, "mupsi_mesh_inside: no such mesh");
  if (n <= 0) return MUPSI_OK;
  if (!px || !py || !pz || !inside)
    return fail(MUPSI_ERR_INVALID, "mupsi_mesh_inside: null buffer");
  return mesh_query_launch(sc, [&](hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_inside, dim3(grid_for(n)), dim3(256), 0, st, *m, px, py, pz, n,
                       inside);
  });
}

extern "C" mupsi_status mupsi_mesh_query_cost(mupsi_scene* sc, int32_t prim, int32_t kind,
                                              const float* px, const float* py, const float* pz,
                                              int64_t n, uint32_t* nodes, uint32_t* tris) {
  const DevMesh* m = pick_mesh(sc, prim);
  if (!m) return fail(MUPSI_ERR_INVALID, "mupsi_mesh_query_cost: no such mesh");
  if (kind < 0 || kind > 4) return fail(MUPSI_ERR_INVALID, "mupsi_mesh_query_cost: kind must be 0 to 4");
  if (n <= 0) return MUPSI_OK;
  if (!px || !py || !pz || !nodes || !tris)
    return fail(MUPSI_ERR_INVALID, "mupsi_mesh_query_cost: null buffer");
  return mesh_query_launch(sc, [&](hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_cost, dim3(query_grid(sc->ctx, n)), dim3(256), 0, st, *m, (int)kind, px, py,
                       pz, n, nodes, tris);
  });
}

extern "C" mupsi_status mupsi_bvh_layout(const float* vertices, int32_t n_vertices,
                                         const int32_t* faces, int32_t n_faces,
                                         int32_t* n_records, int32_t* depth,
                                         int32_t* stackless) {
  if (!vertices || !faces || n_faces < 0 || n_vertices < 0)
    return fail(MUPSI_ERR_INVALID, "mupsi_bvh_layout: bad argument");
  for (int64_t i = 0; i < 3LL * n_faces; ++i)
    if (faces[i] < 0 || faces[i] >= n_vertices)
      return fail(MUPSI_ERR_INVALID, "mupsi_bvh_layout: face index out of range");
  const BvhResult b = build_bvh(vertices, faces, (uint32_t)n_faces);
  BvhRecords L;
  if (!build_records(b, bvh_order(b), L))
    return fail(MUPSI_ERR_INVALID, "mupsi_bvh_layout: the BVH does not fit the record layout");
  if (n_records) *n_records = (int32_t)L.recs.size();
  if (depth) *depth = L.depth;
  if (stackless) *stackless = L.depth <= dev::kStack - 4 ? 1 : 0;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_bvh_xbins(const float* vertices, int32_t n_vertices,
                                        const int32_t* faces, int32_t n_faces, int32_t* ny,
                                        int32_t* nz, float* org_inv, int64_t* n_entries,
                                        float* entries_out, int64_t entries_cap,
                                        uint32_t* offsets_out, int64_t offsets_cap) {
  if (!vertices || !faces || n_faces < 0 || n_vertices < 0 || !ny || !nz || !n_entries)
    return fail(MUPSI_ERR_INVALID, "mupsi_bvh_xbins: bad argument");
  for (int64_t i = 0; i < 3LL * n_faces; ++i)
    if (faces[i] < 0 || faces[i] >= n_vertices)
      return fail(MUPSI_ERR_INVALID, "mupsi_bvh_xbins: face index out of range");
  const BvhResult b = build_bvh(vertices, faces, (uint32_t)n_faces);
  BvhRecords L;
  if (!build_records(b, bvh_order(b), L))
    return fail(MUPSI_ERR_INVALID, "mupsi_bvh_xbins: the BVH does not fit the record layout");
  *ny = *nz = 0;
  *n_entries = 0;
  if (L.depth > dev::kStack - 4 || b.nodes.empty()) return MUPSI_OK;
  const BvhNode& root = b.nodes[b.root];
  XBins xb;
  if (!build_xbins(L, root.bmin, root.bmax, xb)) return MUPSI_OK;
  *ny = xb.ny;
  *nz = xb.nz;
  *n_entries = (int64_t)xb.entries.size();
  if (org_inv) {
    org_inv[0] = xb.y0;
    org_inv[1] = xb.z0;
    org_inv[2] = xb.inv;
  }
  if (entries_out && entries_cap >= (int64_t)xb.entries.size())
    std::memcpy(entries_out, xb.entries.data(), xb.entries.size() * sizeof(XBinEntry));
  if (offsets_out && offsets_cap >= (int64_t)xb.offsets.size())
    std::memcpy(offsets_out, xb.offsets.data(), xb.offsets.size() * sizeof(uint32_t));
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_bvh_build(const float* vertices, int32_t n_vertices,
                                        const int32_t* faces, int32_t n_faces, float* nodes_out,
                                        int32_t* n_nodes_out, uint32_t* tri_index_out,
                                        uint32_t* root_out) {
  if (!vertices || !faces || n_faces < 0 || n_vertices < 0)
    return fail(MUPSI_ERR_INVALID, "mupsi_bvh_build: bad argument");
  for (int64_t i = 0; i < 3LL * n_faces; ++i)
    if (faces[i] < 0 || faces[i] >= n_vertices)
      return fail(MUPSI_ERR_INVALID, "mupsi_bvh_build: face index out of range");
  BvhResult b = build_bvh(vertices, faces, (uint32_t)n_faces);
  if (nodes_out) std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(BvhNode));
  if (n_nodes_out) *n_nodes_out = (int32_t)b.nodes.size();
  if (tri_index_out) std::memcpy(tri_index_out, b.tri_index.data(), b.tri_index.size() * 4);
  if (root_out) *root_out = b.root;
  return MUPSI_OK;
}
