/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the mupsi reference hot path (Eigen-free C++), used by
 * tests/, the smoke check and the CPU-baseline leg of bench.py as the checker.
 * Nothing in the product library links, loads or calls this code.
 *
 * Scene descriptions reuse the plain data structs of include/mupsi.h (the
 * boundary format); all algorithms are restated here independently of the
 * HIP implementation.  Array arguments are host memory.
 *
 * Pinning status of each function is listed in oracle.cpp's header comment.
 */
#ifndef MUPSI_ORACLE_H_
#define MUPSI_ORACLE_H_

#include <stdint.h>
#include "../include/mupsi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- integer RNG core (random.h:13-119, sampler.h:30-61) ---- */
uint32_t orc_make_seed(const uint32_t* vals, int32_t n);
/* n consecutive Random::nextI() outputs.  ctor != 0: Random(seed) constructor
 * semantics (0 -> 1, two warm-up draws); ctor == 0: setState(seed). */
void orc_pcg_draws(uint64_t seed, int32_t ctor, int32_t n, uint32_t* out);
float orc_normalized_uint(uint32_t i);
uint32_t orc_constant_sampler_i(uint32_t seed);

/* ---- GP noise (gpnoise.cpp:51-87, kernel.h:39-78) ---- */
void orc_cell_impulses(const mupsi_gp_params* gp, const int32_t* cell3, const uint32_t* seed,
                       int64_t n, uint32_t* cell_seed, float* pos, int8_t* w);
/* p3: n x 3 (AoS).  fp64 = 0: FP32 faithful sequence; 1: FP64 evaluation with the
 * same integer decisions and FP32 impulse positions.  out4: n x 4, l1: n (may be NULL). */
void orc_psi_eval(const mupsi_gp_params* gp, const float* p3, const uint32_t* seed, int64_t n,
                  int32_t fp64, double* out4, double* l1);
/* Variance normalisation Var() as float (gpnoise.h:31). */
float orc_psi_var(const mupsi_gp_params* gp);

/* ---- scene-level ---- */
void* orc_scene_create(const mupsi_scene_desc* desc);
void orc_scene_free(void* scene);
/* hash11 of the external GP medium (externalgpmedium.cpp:8-17), p3: n x 3 */
void orc_hash11(const float* p3, const uint32_t* seed, int64_t n, float* out);

/* cond: n x 8 floats {active, Cx, Cy, Cz, u_tilde, gx, gy, gz} or NULL (inactive). */
void orc_field_eval(void* scene, const float* p3, const uint32_t* seed, const float* cond,
                    int64_t n, int32_t fp64, double* out4);

/* rays: n x 8 floats {ox, oy, oz, dx, dy, dz, near_t, far_t}.  mode 0 sampleDistance,
 * 1 transmittance.  fp64 = 1 evaluates the field in FP64 at the same FP32 t-grid points.
 * amb[i]: smallest decision margin seen (|f| / (1 + |mu| + l1)); cell[i]: smallest
 * relative distance of a visited point to a cell boundary.  Either may be NULL. */
void orc_march(void* scene, int32_t mode, const float* rays, const uint32_t* seed,
               const float* cond, int64_t n, int32_t fp64, uint8_t* hit, float* t, float* p3,
               float* n3, int32_t* steps, double* amb, double* cell);

/* sampleCondition: u (n), g (n x 3).  fp64 = 0 reference FP32 arithmetic, 1 FP64. */
void orc_condition(void* scene, const float* c3, const float* n3, const uint32_t* seed_next,
                   int64_t n, int32_t fp64, double* u, double* g3);

/* Mean function mu, grad mu at p (meanfunction.cpp:13-39). out4: n x 4. */
void orc_mean_eval(void* scene, const float* p3, int64_t n, double* out4);

/* ---- BVH / mesh (bvh.cpp, mesh.cpp) ---- */
int32_t orc_bvh_build(const float* vertices, int32_t n_vertices, const int32_t* faces,
                      int32_t n_faces, float* nodes_out, uint32_t* tri_index_out,
                      uint32_t* root_out);
/* prim >= 0: primitive index (mesh); prim = -1: the GP mean mesh. */
void orc_mesh_intersect(void* scene, int32_t prim, const float* rays, int64_t n, int32_t* tri,
                        float* t, float* u, float* v);
void orc_mesh_intersect_brute(void* scene, int32_t prim, const float* rays, int64_t n,
                              int32_t* tri, float* t);
void orc_mesh_occluded(void* scene, int32_t prim, const float* rays, int64_t n, uint8_t* occ);
void orc_mesh_closest(void* scene, int32_t prim, const float* p3, int64_t n, float* dist,
                      int32_t* tri, float* closest3);
void orc_mesh_inside(void* scene, int32_t prim, const float* p3, int64_t n, uint8_t* inside);

/* ---- integrator ---- */
/* Camera::generateRay for pixels: out n x 3 directions, origin is cam_pos. */
void orc_camera_rays(void* scene, const int32_t* x, const int32_t* y, int64_t n, float* dir3);
/* Skydrome lookup for directions: rgb n x 3, texel index n. */
void orc_sky_lookup(void* scene, const float* dir3, int64_t n, float* rgb, int32_t* texel);
/* PathTracer::trace for each (x, y, k), seed 42. */
void orc_trace_paths(void* scene, const int32_t* x, const int32_t* y, const int32_t* k,
                     int64_t n, float* rgb, int32_t nthreads);
/* orc_trace_paths plus, per path, the smallest decision margin of its GP-medium
 * marches (each march redone in FP64 at the same FP32 t grid: the smallest
 * |f| / (1 + |mu| + l1) at a visited grid point, 0 where the FP64 decision
 * differs from the FP32 one; +inf for a path with no such march). */
void orc_trace_paths_amb(void* scene, const int32_t* x, const int32_t* y, const int32_t* k,
                         int64_t n, float* rgb, double* amb, int32_t nthreads);
/* Scene::chooseLight at points p3 (n x 3) with the PCG states state[i]
 * (setState semantics; updated in place to the state after the draws).
 * light[i]: primitive index of the light sampled; -1 when no light was drawn
 * (no emitter: no draw at all), -2 when a light was drawn but its sampleDirect
 * failed.  w3 = Le / pdf, d3 = direction, dist, pdf as LightSample. */
void orc_choose_light(void* scene, const float* p3, uint64_t* state, int64_t n, int32_t* light,
                      float* w3, float* d3, float* dist, float* pdf);
/* Renderer::startRenderProgressive restated with OpenMP collapse(2), static
 * schedule, over passes k_begin..k_begin+k_count-1 and the pixels whose linear
 * index is congruent to pixel_offset mod pixel_stride.  Adds radiance into
 * sums (W*H*3, double).  Returns the wall time in seconds. */
double orc_render(void* scene, int32_t k_begin, int32_t k_count, int32_t pixel_stride,
                  int32_t pixel_offset, int32_t nthreads, double* sums, int64_t* paths_out);
int32_t orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
