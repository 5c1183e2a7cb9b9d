/*
 * mupsi.h — C ABI of the MI355X (gfx950) GPIS path tracer.
 *
 * This is the drop-in boundary for the CPU trace loop of mupsi.  The host
 * (C++ scene loader, Python test harness, benchmark driver) describes a scene
 * with the plain structs below, uploads it once, and then asks for spp passes.
 * Every hot loop of the reference trace runs on the GPU behind these calls.
 *
 * Reference interfaces replaced (paths relative to the mupsi source tree):
 *   mupsi_render_passes   <- Renderer::startRenderProgressive, the per-pass
 *                            OpenMP pixel loop body   src/rendering/renderer.cpp:22-49
 *   mupsi_trace_paths     <- PathTracer::trace(pixel, ctx, seed, k)
 *                                                     src/rendering/trace.cpp:118-213
 *   mupsi_read_accum      <- Framebuffer::accumulate / Renderer::getFramebuffer
 *                                                     src/rendering/framebuffer.cpp:28-35,
 *                                                     src/rendering/renderer.h:31
 *   mupsi_scene_upload    <- Controller::load object graph + Renderer::prepareRender
 *                            (Mesh::prepareForRender -> BVH::build)
 *                                                     src/controller/controller.cpp:40-264,
 *                                                     src/rendering/renderer.cpp:10-16
 *   mupsi_psi_eval        <- SparseGPNoiseGenerator::RawNoise   src/gp/gpnoise.h:44,
 *                                                     src/gp/gpnoise.cpp:51-87
 *   mupsi_cell_impulses   <- the per-cell impulse loop of InternalNoise
 *                                                     src/gp/gpnoise.cpp:57-74
 *   mupsi_field_eval      <- GPMedium::evalWithGradient(p, seed, state)
 *                                                     src/medium/gpmedium.cpp:19-41
 *   mupsi_march           <- GPMedium::sampleDistance / GPMedium::transmittance
 *                                                     src/medium/gpmedium.cpp:46-129
 *   mupsi_condition       <- GPMedium::sampleCondition  src/medium/gpmedium.cpp:133-162
 *   mupsi_mesh_intersect  <- Mesh::intersect -> BVH::intersect
 *                                                     src/primitives/mesh.cpp:185-209,
 *                                                     src/bvh/bvh.cpp:197-250
 *   mupsi_mesh_occluded   <- Mesh::occluded -> BVH::occluded  src/primitives/mesh.cpp:236-250,
 *                                                     src/bvh/bvh.cpp:254-300
 *   mupsi_mesh_closest    <- Mesh::distToPoint -> BVH::closest src/primitives/mesh.cpp:325-339,
 *                                                     src/bvh/bvh.cpp:302-353
 *   mupsi_mesh_inside     <- Mesh::inside                src/primitives/mesh.cpp:341-362
 *   mupsi_bvh_build       <- BVH::build / buildRecursive src/bvh/bvh.cpp:115-193
 *   mupsi_scene_load_json <- Controller::load (scene.json schema)
 *                                                     src/controller/controller.cpp:40-264
 *   mupsi_scene_set_gp_medium <- Controller::applyGpMedium  src/controller/controller.cpp:397-426
 *   mupsi_scene_set_settings  <- the RenderContext snapshot of Controller::startProgressive
 *                                                     src/controller/controller.cpp:293-345
 *
 * Conventions
 *   - Every function returns mupsi_status (0 = OK).  No C++ exception crosses
 *     the ABI.  mupsi_last_error() returns a message for the last failure on
 *     the calling thread.
 *   - Scene descriptors are host memory owned by the caller; upload copies, so
 *     the caller may free them on return.
 *   - Batch entry points (psi_eval, march, mesh_* ...) take DEVICE pointers
 *     (HBM-resident SoA arrays) and run synchronously on the context stream.
 *   - One HIP stream per context; a context is not thread-safe.  Distinct
 *     contexts may be used from distinct threads.
 *   - The library is GPU-only: without a usable gfx950 device mupsi_open
 *     fails with MUPSI_ERR_HIP.  There is no CPU fallback.
 */
#ifndef MUPSI_H_
#define MUPSI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mupsi_status;
enum {
  MUPSI_OK = 0,
  MUPSI_ERR_INVALID = 1,     /* bad argument / descriptor */
  MUPSI_ERR_HIP = 2,         /* HIP runtime failure or no device */
  MUPSI_ERR_NOMEM = 3,       /* device or host allocation failed */
  MUPSI_ERR_ABORTED = 4,     /* abort flag observed; accumulation kept at last full pass */
  MUPSI_ERR_UNSUPPORTED = 5, /* descriptor uses a feature this build does not lower */
  MUPSI_ERR_IO = 6,          /* file could not be read */
  MUPSI_ERR_PARSE = 7        /* malformed scene.json / OBJ / image */
};

typedef struct mupsi_ctx mupsi_ctx;
typedef struct mupsi_scene mupsi_scene;

/* ---- scene description (mirrors the scene.json schema, controller.cpp:40-264) ---- */

enum { MUPSI_GP_SINGLE_REALIZATION = 0, MUPSI_GP_RENEWAL_PLUS = 1 };
enum { MUPSI_MEAN_SPHERE = 0, MUPSI_MEAN_MESH = 1 };
enum { MUPSI_PRIM_SPHERE = 0, MUPSI_PRIM_MESH = 1 };
enum { MUPSI_BSDF_LAMBERTIAN = 0, MUPSI_BSDF_NULL = 1, MUPSI_BSDF_SPECULAR = 2 };
/* The medium inside int_medium spheres (controller.cpp:80-190, 230-234: gp_medium
 * wins over sdf_medium, which wins over external_gp_medium). */
enum { MUPSI_MEDIUM_GP = 0, MUPSI_MEDIUM_SDF = 1, MUPSI_MEDIUM_GRID = 2 };

/* SparseSEKernel(sigma, kernelRadius = kernel.length_scale, lengthScale = aniso)
 * + SparseGPNoiseGenerator(kernel, points_per_cell)   controller.cpp:112-120 */
typedef struct {
  float sigma;           /* kernel.sigma */
  float kernel_radius;   /* kernel.length_scale: the impulse cell size kr */
  float aniso[3];        /* kernel.aniso: SE scale L^2 = |aniso|^2 */
  int32_t points_per_cell;
} mupsi_gp_params;

/* Triangle mesh as parsed from OBJ (Mesh::fetchFrom, mesh.cpp:20-136) placed by
 * Translation(position) * Scaling(scale) (controller.cpp:102-105,249-253). */
typedef struct {
  const float* vertices;    /* n_vertices x 3 */
  const float* normals;     /* n_normals x 3, indexed by VERTEX index, or NULL */
  const float* texcoords;   /* n_texcoords x 2, or NULL */
  const int32_t* faces;     /* n_faces x 3 vertex indices (0-based) */
  const int32_t* face_tex;  /* n_faces x 3 texcoord indices, or NULL */
  int32_t n_vertices, n_normals, n_texcoords, n_faces;
  float position[3];
  float scale;
} mupsi_mesh_desc;

/* ExternalGPMedium (externalgpmedium.h:10-38): mean and optional variance
 * GridMeanFunction grids (meanfunction.cpp:41-86), res^3 floats each, x-major
 * (index (i * res + j) * res + k), over the box [bounds_min, bounds_max]. */
typedef struct {
  const float* mean;
  const float* variance;   /* NULL: deterministic medium */
  int32_t resolution;      /* external_gp_medium.resolution (48) */
  float bounds_min[3], bounds_max[3];
  float noise_scale;       /* external_gp_medium.noise_scale (1) */
} mupsi_grid_medium_desc;

/* RGB float texture, row-major, row 0 at the top (BitmapTexture, texture.cpp:52-69). */
typedef struct {
  const float* rgb;  /* w x h x 3, or NULL for "absent" */
  int32_t width, height;
} mupsi_texture_desc;

typedef struct {
  int32_t type;            /* MUPSI_PRIM_* */
  int32_t bsdf;            /* MUPSI_BSDF_* */
  /* sphere */
  float center[3];
  float radius;
  int32_t has_emission;    /* sphere "emission" key */
  float emission[3];
  int32_t int_medium;      /* sphere "int_medium" key present: GP medium inside */
  /* mesh */
  mupsi_mesh_desc mesh;
  mupsi_texture_desc albedo; /* lambertian mesh "texture"; rgb NULL = constant 0.8 */
} mupsi_primitive_desc;

typedef struct {
  int32_t width, height;
  float cam_pos[3], cam_dir[3], cam_up[3];
  float fov;                   /* degrees in the schema; the reference feeds it to tan() as radians */
  int32_t max_bounce;          /* PathTracerSettings::max_bounce */
  int32_t gp_mode;             /* MUPSI_GP_* */
  int32_t skip_medium;         /* PathTracerSettings::skip_medium */
  int32_t has_gp_medium;
  int32_t mean_type;           /* MUPSI_MEAN_* */
  float mean_center[3];
  float mean_radius;
  mupsi_mesh_desc mean_mesh;
  mupsi_gp_params kernel;
  mupsi_texture_desc sky;      /* skydrome texture; rgb NULL = no skydrome */
  const mupsi_primitive_desc* primitives;
  int32_t n_primitives;
  /* which medium has_gp_medium describes: MUPSI_MEDIUM_GP (mean_* + kernel),
   * MUPSI_MEDIUM_SDF (SDFMedium over the mean_* function, sdfmedium.cpp:16-98)
   * or MUPSI_MEDIUM_GRID (ExternalGPMedium over grid, externalgpmedium.cpp:9-130) */
  int32_t medium_kind;
  mupsi_grid_medium_desc grid;
} mupsi_scene_desc;

/* Conditioning state (GPConditioningState, medium.h:12-17) as SoA device arrays. */
typedef struct {
  const uint8_t* active;   /* n */
  const float* cx; const float* cy; const float* cz;
  const float* u_tilde;
  const float* gx; const float* gy; const float* gz;
} mupsi_cond_soa;

/* ---- context ---- */
mupsi_status mupsi_open(int32_t device, mupsi_ctx** out);
/* Frees the context and every scene still open on it; those scene handles become invalid. */
void mupsi_close(mupsi_ctx* ctx);
const char* mupsi_status_string(mupsi_status s);
const char* mupsi_last_error(void);
/* Device name and compute-unit count of the context's GPU. */
mupsi_status mupsi_device_info(mupsi_ctx* ctx, char* name, int32_t name_len, int32_t* cu_count);

/* ---- scene ---- */
mupsi_status mupsi_scene_upload(mupsi_ctx* ctx, const mupsi_scene_desc* desc, mupsi_scene** out);
/* Parse a scene.json file with the reference schema and upload it.  spp_out
 * receives the "spp" key (1 if absent), as Controller::load stores spp_. */
mupsi_status mupsi_scene_load_json(mupsi_ctx* ctx, const char* path, mupsi_scene** out, int32_t* spp_out);
void mupsi_scene_free(mupsi_scene* scene);
mupsi_status mupsi_scene_size(mupsi_scene* scene, int32_t* width, int32_t* height);
/* Controller::applyGpMedium (controller.cpp:397-426): give every int_medium
 * primitive a new GPMedium over the mean (MUPSI_MEAN_SPHERE: mean_center,
 * mean_radius; MUPSI_MEAN_MESH: mean_mesh) with SparseSEKernel(kernel).  Only
 * for scenes whose medium is a gp_medium; the sums are kept (reset them as
 * Renderer::prepareRender does); refused while the scene is in a live batch. */
mupsi_status mupsi_scene_set_gp_medium(mupsi_scene* scene, int32_t mean_type,
                                       const float* mean_center, float mean_radius,
                                       const mupsi_mesh_desc* mean_mesh,
                                       const mupsi_gp_params* kernel);
/* The RenderContext snapshot (render_context.h:15-19) of g_pathTracerSettings
 * and g_gpSettings: max_bounce, gp_mode (MUPSI_GP_*) and skip_medium, as
 * Controller::setGpMode (controller.cpp:389-395) and
 * Controller::startProgressive (controller.cpp:282-289) change them.  Refused
 * while the scene is in a live batch. */
mupsi_status mupsi_scene_set_settings(mupsi_scene* scene, int32_t max_bounce, int32_t gp_mode,
                                      int32_t skip_medium);

/* Render spp passes k_begin .. k_begin+k_count-1 of every pixel and add the
 * path radiance to the FP64 per-pixel sums.  abort_flag (host memory, may be
 * NULL) is polled between bounce iterations and passes; on abort the sums are
 * restored to the last completed call and MUPSI_ERR_ABORTED is returned.
 * passes_per_launch bounds how many passes run as one wavefront (0 = auto). */
mupsi_status mupsi_render_passes(mupsi_scene* scene, int32_t k_begin, int32_t k_count,
                                 int32_t passes_per_launch, const volatile int32_t* abort_flag);
/* No reference counterpart (how the pass loop is launched): with use_graph != 0,
 * a plain render (no guide) called without an abort flag captures the bounce
 * loop of a launch (init, marches, shades) into a hipGraph once and replays it
 * for every later launch of the same size, the first pass read from device
 * memory; the results are those of the launches issued one by one.  A capture
 * is dropped when the path buffers grow or the scene's settings or GP medium
 * change.  mupsi_scene_graph_builds counts the captures made. */
mupsi_status mupsi_scene_set_graph(mupsi_scene* scene, int32_t use_graph);
mupsi_status mupsi_scene_graph_builds(mupsi_scene* scene, int64_t* builds);
/* No reference counterpart (where the shading of a medium hit runs): with
 * fused != 0 (the default) the GP-medium marcher of a render shades the hits of
 * a sphere-mean GP medium (at most 8 points per cell) itself --
 * PathTracer::handleVolume without emitters
 * (trace.cpp:77-116) and GPMedium::sampleCondition (gpmedium.cpp:133-162) --
 * and marches the path's next ray in the same launch, when the scene has no
 * emitter and no guide is attached; other scenes are not affected.  The
 * results are bit for bit those of fused = 0, where every hit goes back to the
 * shading kernel of the bounce loop. */
mupsi_status mupsi_scene_set_fused(mupsi_scene* scene, int32_t fused);
mupsi_status mupsi_reset_accum(mupsi_scene* scene);
/* rgb_sum: host, width*height*3 doubles (row-major, y = row); sample_count: passes accumulated. */
mupsi_status mupsi_read_accum(mupsi_scene* scene, double* rgb_sum, int64_t* sample_count);
/* Replace the sums (resume from a checkpoint written from mupsi_read_accum). */
mupsi_status mupsi_write_accum(mupsi_scene* scene, const double* rgb_sum, int64_t sample_count);

/* Per-path replay: radiance of PathTracer::trace((x[i], y[i]), ctx, 42, k[i]).
 * Host arrays of length n; rgb_out host n x 3. */
mupsi_status mupsi_trace_paths(mupsi_scene* scene, const int32_t* x, const int32_t* y,
                               const int32_t* k, int64_t n, float* rgb_out);

/* Work counters of the last mupsi_render_passes call (summed over its launches). */
typedef struct {
  int64_t paths;             /* paths traced */
  int64_t march_calls;       /* sampleDistance invocations */
  int64_t march_steps;       /* t-grid points visited (reference evaluations) */
  int64_t field_evals;       /* field evaluations actually executed */
  int64_t impulse_evals;     /* SE kernel evaluations (field_evals x 27 x ppc) */
  int64_t march_kernel_launches;
  double march_kernel_ms;    /* summed HIP-event time of the march kernel */
  double total_ms;           /* wall time of the call */
} mupsi_stats;
mupsi_status mupsi_last_stats(mupsi_scene* scene, mupsi_stats* out);

/* ---- batch entry points on device pointers ---- */

/* RawNoise(p, seed): out4[i] = (psi, dpsi/dx, dpsi/dy, dpsi/dz).
 * l1 (may be NULL): sigma/sqrt(Var) * sum_i |h_i| per point (error scale). */
mupsi_status mupsi_psi_eval(mupsi_ctx* ctx, const mupsi_gp_params* gp,
                            const float* px, const float* py, const float* pz,
                            const uint32_t* seed, int64_t n, float* out4, float* l1);

/* Impulses of cells cell3[i] = (cx, cy, cz) for realization seed[i]:
 * cell_seed[i] = make_seed(cx, cy, cz, seed); pos[(i*ppc + j)*3 + a]; w[i*ppc + j] = +-1. */
mupsi_status mupsi_cell_impulses(mupsi_ctx* ctx, const mupsi_gp_params* gp,
                                 const int32_t* cell3, const uint32_t* seed, int64_t n,
                                 uint32_t* cell_seed, float* pos, int8_t* w);

/* GPMedium::evalWithGradient(p, seed, state) with the scene's GP medium. */
mupsi_status mupsi_field_eval(mupsi_scene* scene, const float* px, const float* py,
                              const float* pz, const uint32_t* seed,
                              const mupsi_cond_soa* cond, int64_t n, float* out4);

/* GPMedium::sampleDistance (mode 0) or ::transmittance (mode 1) of rays
 * (ox..dz, near_t, far_t; directions are normalized as Ray's constructor does).
 * Outputs: hit[i] (1 = sign change found; for mode 1 this means opaque),
 * t[i], p[i*3], nrm[i*3] (mode 0 only), steps[i] = t-grid points visited. */
mupsi_status mupsi_march(mupsi_scene* scene, int32_t mode,
                         const float* ox, const float* oy, const float* oz,
                         const float* dx, const float* dy, const float* dz,
                         const float* near_t, const float* far_t, const uint32_t* seed,
                         const mupsi_cond_soa* cond, int64_t n,
                         uint8_t* hit, float* t, float* p, float* nrm, int32_t* steps);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mupsi_march with the same arguments and results, whose
 * scratch (the marcher's work counter and statistics, the normalized ray
 * copies and hit points) is a fixed array in the library's device memory,
 * 9 MiB per GPU, held until the process ends, instead of per-call allocations.
 * n above MUPSI_MARCH_NOALLOC_MAX_RAYS is refused with MUPSI_ERR_INVALID.
 * Every call on one GPU uses that GPU's one array: calls that overlap on one
 * GPU (from two threads or two contexts) race. */
enum { MUPSI_MARCH_NOALLOC_MAX_RAYS = 262144 };
mupsi_status mupsi_march_noalloc(mupsi_scene* scene, int32_t mode,
                                 const float* ox, const float* oy, const float* oz,
                                 const float* dx, const float* dy, const float* dz,
                                 const float* near_t, const float* far_t, const uint32_t* seed,
                                 const mupsi_cond_soa* cond, int64_t n,
                                 uint8_t* hit, float* t, float* p, float* nrm, int32_t* steps);

/* GPMedium::sampleCondition for medium hits at C with normal nrm, next-realization
 * seed: the Renewal+ 4x4 system solved in FP64.  u_tilde[n], g_tilde[n*3] (double). */
mupsi_status mupsi_condition(mupsi_scene* scene, const float* cx, const float* cy,
                             const float* cz, const float* nx, const float* ny,
                             const float* nz, const uint32_t* seed_next, int64_t n,
                             double* u_tilde, double* g_tilde);

/* Mesh queries against primitive index prim (MUPSI_PRIM_MESH) or, when
 * prim == -1, the GP mean mesh.  World-space inputs, reference semantics. */
mupsi_status mupsi_mesh_intersect(mupsi_scene* scene, int32_t prim,
                                  const float* ox, const float* oy, const float* oz,
                                  const float* dx, const float* dy, const float* dz,
                                  const float* near_t, const float* far_t, int64_t n,
                                  int32_t* tri, float* t, float* u, float* v);
mupsi_status mupsi_mesh_occluded(mupsi_scene* scene, int32_t prim,
                                 const float* ox, const float* oy, const float* oz,
                                 const float* dx, const float* dy, const float* dz,
                                 const float* near_t, const float* far_t, int64_t n,
                                 uint8_t* occluded);
/* mupsi_mesh_closest on the GP mean mesh (prim -1) starts the search from the
 * mesh's distance-grid bound; distance, triangle and point are those of the
 * unbounded search. */
mupsi_status mupsi_mesh_closest(mupsi_scene* scene, int32_t prim, const float* px,
                                const float* py, const float* pz, int64_t n,
                                float* dist, int32_t* tri, float* closest3);
mupsi_status mupsi_mesh_inside(mupsi_scene* scene, int32_t prim, const float* px,
                               const float* py, const float* pz, int64_t n, uint8_t* inside);

/* Traversal cost of the mesh-mean queries at n points: per point, the BVH
 * nodes whose bounds were tested and the triangles tested by
 * Mesh::distToPoint -> BVH::closest (kind 0; src/primitives/mesh.cpp:325-339,
 * src/bvh/bvh.cpp:302-353, where both children of an expanded node are
 * tested), Mesh::inside -> BVH::intersect (kind 1; mesh.cpp:341-362,
 * bvh.cpp:197-250, every re-cast included), or distToPoint started from the
 * distance-grid bound as mupsi_mesh_closest runs it on a mean mesh (kind 2),
 * or Mesh::inside as mupsi_mesh_inside runs it: one pass collecting the +x
 * ray's hits over the leaves of the point's bin where the mesh carries leaf
 * bins (mean meshes; a node count is a bin entry read), the reference's casts
 * where that cannot decide (kind 3), or the same with the stackless walk over
 * the tree in place of the bins (kind 4).  The
 * reference keeps no such counters; they give the traversal's algorithmic
 * bytes (32 per node, 52 per triangle: index, face and three vertices). */
mupsi_status mupsi_mesh_query_cost(mupsi_scene* scene, int32_t prim, int32_t kind,
                                   const float* px, const float* py, const float* pz,
                                   int64_t n, uint32_t* nodes, uint32_t* tris);

/* Device time in milliseconds (HIP events on the context's stream) of the
 * kernel run by the last mupsi_mesh_* call on this context.  No reference
 * counterpart: it times the mesh-query kernels for the traversal roofline. */
mupsi_status mupsi_last_kernel_ms(mupsi_ctx* ctx, float* ms);

/* Host-side BVH build with the reference topology (post-order node array, the
 * root is the last node).  nodes_out: 8 floats/ints per node packed as
 * {minx,miny,minz,maxx,maxy,maxz, left(u32 bits), right(u32 bits)};
 * capacity 2*n_faces nodes.  tri_index_out: n_faces entries. */
mupsi_status mupsi_bvh_build(const float* vertices, int32_t n_vertices, const int32_t* faces,
                             int32_t n_faces, float* nodes_out, int32_t* n_nodes_out,
                             uint32_t* tri_index_out, uint32_t* root_out);

/* Device layout of the BVH mupsi_bvh_build builds (no reference counterpart:
 * it reports how the library stores and walks the reference's tree): the
 * number of 64-byte records (one per internal node, both children's bounds
 * side by side, breadth-first so that the top levels come first), the levels
 * of internal nodes on the longest root-to-leaf path, and whether
 * Mesh::inside runs as one stackless walk (trees of at most 60 levels, where
 * the reference's 64-entry stack cannot fill) or as the reference's casts. */
mupsi_status mupsi_bvh_layout(const float* vertices, int32_t n_vertices, const int32_t* faces,
                              int32_t n_faces, int32_t* n_records, int32_t* depth,
                              int32_t* stackless);

/* The leaf bins a mean mesh carries for Mesh::inside (no reference counterpart:
 * it reports how the library finds the leaves the +x parity ray of
 * src/primitives/mesh.cpp:341-362 can reach): the local-space y-z box cut into
 * ny x nz square bins of side 1/org_inv[2] (a power of two) from (org_inv[0],
 * org_inv[1]); every leaf listed in each bin its y-z bounds reach, 8 words per
 * entry {xmax, ymin, ymax, zmin, zmax, xmin, leaf reference bits, 0}, sorted by
 * xmax, largest first; bin iz * ny + iy holds entries offsets[b] ..
 * offsets[b + 1] - 1.  ny = nz = 0: no bins (a single leaf, or a tree deeper
 * than the stackless walk allows).  entries_out / offsets_out are written when
 * non-null and at least n_entries / ny * nz + 1 long. */
mupsi_status mupsi_bvh_xbins(const float* vertices, int32_t n_vertices, const int32_t* faces,
                             int32_t n_faces, int32_t* ny, int32_t* nz, float* org_inv,
                             int64_t* n_entries, float* entries_out, int64_t entries_cap,
                             uint32_t* offsets_out, int64_t offsets_cap);

/* ---- batched scenes ----
 * Several uploaded scenes of one context traced in shared wavefront launches
 * (the pass loop of Renderer::startRenderProgressive, renderer.cpp:22-49, for
 * many scenes at once): every launch mixes the paths of several scenes, each
 * path carrying its scene id, so rays of uneven march lengths from all scenes
 * share the persistent marchers.  Each scene accumulates into its own
 * accumulator (mupsi_read_accum), with the same per-path results as
 * mupsi_render_passes.  The scenes must outlive the batch; light samples of
 * medium hits are marched serially in batched launches and attached guides
 * are not used. */
typedef struct mupsi_batch mupsi_batch;

typedef struct {
  int64_t paths_per_launch;    /* 0: chosen by the autotuner on the first launches */
  int32_t march_blocks_per_cu; /* persistent marcher blocks per CU; 0: autotuned */
  int32_t use_graph;           /* 1: each launch's bounce loop replays a captured HIP graph */
} mupsi_batch_tuning;

typedef struct {
  int64_t paths;               /* paths traced by mupsi_batch_render calls */
  int64_t launches;
  double total_ms;             /* wall time of the render calls */
  double tune_ms;              /* part of it spent on autotuner probe launches (real passes) */
  int64_t graph_builds;
} mupsi_batch_stats;

mupsi_status mupsi_batch_create(mupsi_ctx* ctx, mupsi_scene* const* scenes, int32_t n,
                                mupsi_batch** out);
void mupsi_batch_free(mupsi_batch* batch);
mupsi_status mupsi_batch_set_tuning(mupsi_batch* batch, const mupsi_batch_tuning* tuning);
mupsi_status mupsi_batch_get_tuning(mupsi_batch* batch, mupsi_batch_tuning* tuning);
/* Passes k_begin[s] .. k_begin[s] + k_count[s] - 1 of every scene s. */
mupsi_status mupsi_batch_render(mupsi_batch* batch, const int32_t* k_begin, const int32_t* k_count,
                                const volatile int32_t* abort_flag);
mupsi_status mupsi_batch_stats_get(mupsi_batch* batch, mupsi_batch_stats* out);

/* ---- learned path guiding ----
 * No counterpart in the reference: an addition to its trace loop
 * (trace.cpp:77-116, the Lambertian sample of handleVolume).  A guide is a
 * 32 -> 64 -> 64 -> 64 -> 41 MLP (BF16 MFMA, FP32 accumulate, FP32 master
 * parameters, Adam) whose output is an 8-lobe von Mises-Fisher mixture over
 * directions plus the probability alpha of sampling it at this vertex.
 * Attached to a scene, it proposes the direction at every GP-medium hit; the
 * proposal is one-sample MIS-combined with the Lambertian cosine sample, so the
 * image keeps the plain tracer's expectation.  The first train_passes
 * passes rendered with the guide attached train it online on the radiance the
 * tracer itself measures. */
typedef struct mupsi_guide mupsi_guide;

typedef struct {
  uint64_t seed;           /* parameter initialisation and minibatch draws */
  float learning_rate;     /* Adam step size (1e-3) */
  float alpha;             /* largest guide selection probability (0.9; 0 = never) */
  int32_t train_passes;    /* passes of online training (32) */
  int32_t batch_size;      /* training minibatch, multiple of 64 (16384) */
  int32_t min_steps_per_pass; /* Adam steps per training pass: about four epochs of */
  int32_t max_steps_per_pass; /*   the pass's new records, within [min, max] */
  int32_t max_records;     /* guided vertices recorded per path for training (1..4) */
  float box_center[3];     /* position normalisation of the encoding: */
  float box_radius;        /*   p_hat = clamp((p - center) / radius, -1, 1) */
  float indirect_clamp;    /* training radiance: direct + min(indirect, clamp) (luminance) */
  float rr_boost;          /* guided roulette survives with min(1, rr_boost max(T)) (>= 1) */
  int32_t passes_per_update; /* training passes traced per launch before the Adam steps (4) */
} mupsi_guide_config;

typedef struct {
  int64_t train_passes_done;
  int64_t train_steps;
  int64_t train_samples;     /* training records collected */
  double last_loss;          /* mean minibatch loss of the last step */
  double train_ms;           /* time in training kernels (HIP events) */
  double infer_ms;           /* time in inference kernels (HIP events) */
} mupsi_guide_stats;

enum {
  MUPSI_GUIDE_NUM_PARAMS = 13097,
  MUPSI_GUIDE_NUM_INPUTS = 32,
  MUPSI_GUIDE_NUM_OUTPUTS = 41,
  MUPSI_GUIDE_MIXTURE_FLOATS = 44  /* 8 x (mu.x, mu.y, mu.z, kappa, weight), alpha, 3 unused */
};

mupsi_status mupsi_guide_create(mupsi_ctx* ctx, const mupsi_guide_config* cfg, mupsi_guide** out);
void mupsi_guide_free(mupsi_guide* guide);
/* FP32 parameters, layout W1[64][32] b1[64] W2[64][64] b2 W3[64][64] b3 W4[41][64] b4[41]. */
mupsi_status mupsi_guide_get_params(mupsi_guide* guide, float* host_params);
/* Replaces the parameters and resets Adam's moments and step count. */
mupsi_status mupsi_guide_set_params(mupsi_guide* guide, const float* host_params);
mupsi_status mupsi_guide_stats_get(mupsi_guide* guide, mupsi_guide_stats* out);
/* Attach (or, with guide = NULL, detach) a guide; resets its training pass count. */
mupsi_status mupsi_scene_set_guide(mupsi_scene* scene, mupsi_guide* guide);

/* Training samples the tracer recorded for the attached guide (the first
 * min(capacity, stored) records of its ring): features n x 32, sampled
 * directions n x 3, weights lum(Li) max(cos, 0) / q; host arrays. */
mupsi_status mupsi_scene_guide_records(mupsi_scene* scene, float* feats, float* dirs,
                                       float* weights, int64_t capacity, int64_t* count);

/* Batch entry points on device pointers (inference, gradients, one Adam step). */
/* feats n x 32 f32 -> raw n x 41 outputs and/or mixture n x 44. */
mupsi_status mupsi_guide_forward(mupsi_guide* guide, const float* feats, int64_t n, float* raw,
                                 float* mixture);
/* Loss sum_i -w_i log q(dir_i | feats_i) / n of the combined density
 * q = alpha p_guide + (1 - alpha) max(cos, 0)/pi (cos against the normal in
 * feats[24..26]) and its gradient (host, 13097 floats). */
mupsi_status mupsi_guide_gradient(mupsi_guide* guide, const float* feats, const float* dirs,
                                  const float* weights, int64_t n, float* grad_host,
                                  double* loss_host);
/* One Adam step on the minibatch (all n samples, n <= 65536). */
mupsi_status mupsi_guide_train_step(mupsi_guide* guide, const float* feats, const float* dirs,
                                    const float* weights, int64_t n);
/* Encoding of (p, n, wo), each n x 3, into n x 32 features. */
mupsi_status mupsi_guide_encode(mupsi_guide* guide, const float* p, const float* nrm,
                                const float* wo, int64_t n, float* feats);
/* The tracer's guided sample at a Lambertian vertex: u n x 4 uniforms (technique,
 * lobe, cosine, azimuth; the cosine branch uses u[1], u[2] as its r1, r2),
 * mixture n x 44, normals n x 3 -> dir n x 3, combined pdf q n, weight n
 * (albedo-free f cos / q). */
mupsi_status mupsi_guide_sample(mupsi_guide* guide, const float* mixture, const float* nrm,
                                const float* u, int64_t n, float* dir, float* pdf,
                                float* weight);

#ifdef __cplusplus
}
#endif
#endif /* MUPSI_H_ */
