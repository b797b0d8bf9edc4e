/*
 * psdr_hip.h — C ABI of the MI355X (gfx950) differentiable path tracer.
 *
 * This is the drop-in boundary for torch_psdr's hot path.  The reference binds
 * its C++ classes to Python with pybind11 (module `pypsdr`, src/psdr.cc:15-137);
 * these entry points replace that module for the render path.  All buffers are
 * device pointers owned by the caller (torch tensors); the library owns only what
 * it derives from them (triangle records, edges, samplers, LBVH, wavefront
 * queues) inside an opaque scene handle.  Every call is asynchronous on the
 * given HIP stream (NULL = the null stream) and never synchronises the host
 * unless stated.  Every call returns 0 on success or a negative code; the text
 * of the last error on the calling thread is returned by psdr_last_error().
 * No C++ exception crosses this boundary.
 *
 * Graph capture (no reference counterpart; the reference synchronises after
 * every ray cast, src/scene_optix.cc:275): psdr_scene_update,
 * psdr_render_forward and psdr_render_backward may be recorded into a HIP graph
 * by stream capture of `stream` (hipStreamBeginCapture, torch.cuda.graph) once
 * the same calls, with the same sizes, have run eagerly on the scene: they then
 * allocate nothing and never synchronise the host, and the LBVH rebuild of an
 * update is part of the graph.  The boundary terms of a renderD run on two
 * auxiliary streams forked from and joined back into the captured stream by the
 * forward / backward pair, so a capture must hold both halves.  A call that
 * would have to grow a scene buffer during capture fails with
 * PSDR_ERR_CAPTURE; once any capture has happened in the process, a buffer that
 * grows keeps its old allocation until the process ends (a graph may point at
 * it).  Replays and eager calls share the scene's buffers: they must not run
 * concurrently on one scene.  The *_plain entry points synchronise the host and
 * fail with PSDR_ERR_CAPTURE under capture.
 */
#ifndef PSDR_HIP_H
#define PSDR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSDR_ABI_VERSION 7

/* status codes */
#define PSDR_OK 0
#define PSDR_ERR_ARG -1      /* bad argument (shape, size, null pointer) */
#define PSDR_ERR_HIP -2      /* a HIP runtime call failed */
#define PSDR_ERR_STATE -3    /* call out of order (e.g. backward without forward) */
#define PSDR_ERR_NOMEM -4    /* device allocation failed */
#define PSDR_ERR_CAPTURE -5  /* not possible under stream capture (see above) */

/* Material kinds (src/material.h:24-58). */
#define PSDR_MAT_DIFFUSE 0      /* DiffuseMaterial: albedo texture / pi      */
#define PSDR_MAT_DIFFUSE_BSDF 1 /* DiffuseBsdfMaterial: color, roughness     */
#define PSDR_MAT_LIGHT 2        /* DiffuseLightMaterial of an AreaLight mesh */
/* The bounding cube of an environment light (src/emitter.cc:95-130, 218-247;
 * placed by Scene::Scene, src/scene.cc:31-55): 12 triangles, the LAST object,
 * at most one.  Its radiance is the cube map at the hit's uv (EnvLight,
 * albedo_* = the map) or the second-order spherical harmonics of the hit's
 * direction to the viewer (SH9Light, albedo_offset = 27 coefficient floats, 9
 * RGB rows in the texel buffer).  It has no edges, its light power is 4*pi and
 * the traversal tests its triangles for every ray instead of holding them in
 * the LBVH. */
#define PSDR_MAT_ENV_MAP 3
#define PSDR_MAT_ENV_SH9 4

/* One object = one mesh (scene objects are meshes first, then the meshes of area
 * lights, as in Scene::Scene, src/scene.cc:8-20).  Textures live in one packed
 * texel buffer; a texture with h == w == 0 is constant (ConstantTexture). */
typedef struct psdr_object {
  int32_t material;        /* PSDR_MAT_* */
  int32_t light;           /* index into the light arrays, or -1 */
  int32_t albedo_offset;   /* float offset of the 3-channel albedo / color texture */
  int32_t albedo_h, albedo_w;
  int32_t rough_offset;    /* float offset of the 1-channel roughness texture */
  int32_t rough_h, rough_w;
  int32_t tri_begin, tri_count; /* triangle range of this object */
  int32_t grad_mask;       /* PSDR_GRAD_* bits: which of this object's inputs need gradients */
  int32_t normal_offset;   /* float offset of the 3-channel normal map (DiffuseBsdf), or -1 for none */
  int32_t normal_h, normal_w;
} psdr_object;

/* grad_mask bits (mirror `requires_grad` of the wrapped tensors,
 * src/python/torch_psdr.py:84-103) */
#define PSDR_GRAD_VERTICES 1
#define PSDR_GRAD_ALBEDO 2
#define PSDR_GRAD_ROUGHNESS 4
#define PSDR_GRAD_EMIT 8
#define PSDR_GRAD_NORMAL 16

/* Scene description.  Device arrays: vertices, faces, tri_uv, tri_obj, texels,
 * light_emit.  Host arrays: objects, light_obj. */
typedef struct psdr_scene_desc {
  int32_t n_vertices;
  const float* vertices;   /* [n_vertices*3]                           */
  int32_t n_tris;
  const int32_t* faces;    /* [n_tris*3] global vertex ids             */
  const float* tri_uv;     /* [n_tris*6] corner uvs                    */
  const int32_t* tri_obj;  /* [n_tris]                                 */
  int32_t n_objects;
  const psdr_object* objects; /* host [n_objects]                     */
  int32_t n_texels;        /* floats in texels                          */
  const float* texels;
  int32_t n_lights;
  const float* light_emit; /* [n_lights*3]                              */
  const int32_t* light_obj;/* host [n_lights]                           */
} psdr_scene_desc;

/* PathIntegrator::Config (src/path_integrator.h:8-21; pybind fields
 * src/psdr.cc:108-131).  Defaults as in the Python wrapper
 * (src/python/torch_psdr.py:478-491). */
typedef struct psdr_config {
  uint32_t n_pass;
  uint32_t spp_interior;
  uint32_t enable_light_visable;
  uint32_t spp_primary_edge;
  uint32_t spp_secondary_edge;
  uint32_t max_bounce;
  uint32_t mis_light_samples;
  uint32_t mis_bsdf_samples;
  uint32_t primary_edge_preprocess_rounds;
  uint32_t secondary_edge_preprocess_rounds;
  uint32_t primary_edge_hypercube_resolution;
  uint32_t secondary_edge_hypercube_resolution[3];
} psdr_config;

/* Camera kinds: PerspectiveCamera(look_from, look_at, up, vfov[rad], height,
 * width) (src/psdr.cc:22-29, src/camera.h:208-277) and
 * OrthographicCamera(look_from, look_at, up, view_width, height, width)
 * (src/psdr.cc:30-37, src/camera.h:279-362). */
#define PSDR_CAMERA_PERSPECTIVE 0
#define PSDR_CAMERA_ORTHOGRAPHIC 1

typedef struct psdr_camera {
  float look_from[3];
  float look_at[3];
  float up[3];
  float vfov;        /* perspective: vertical field of view in radians */
  int32_t height;
  int32_t width;
  int32_t kind;      /* PSDR_CAMERA_* */
  float view_width;  /* orthographic: visible width in world units */
} psdr_camera;

/* Camera gradient layout (PSDR_CAMERA_GRAD floats): d/d look_from[3],
 * look_at[3], up[3], then vfov (perspective) or view_width (orthographic). */
#define PSDR_CAMERA_GRAD 10

typedef struct psdr_scene psdr_scene; /* opaque */

/* ---- errors / version ---------------------------------------------------- */
const char* psdr_last_error(void);
int psdr_abi_version(void);

/* ---- scene ---------------------------------------------------------------- */
/* Most triangles a scene may hold: triangle records are indexed as 5 * t in
 * 32-bit arithmetic and LBVH nodes carry child ids in 30 bits.  A larger
 * n_tris is refused with PSDR_ERR_ARG. */
#define PSDR_MAX_TRIS 268435456

/* Replaces pypsdr.TriangleMesh / AreaLight / Scene construction
 * (src/psdr.cc:79-106 -> src/mesh.cc:5-95, src/emitter.cc:4-14, src/scene.cc:3-104,
 * src/scene_optix.cc:10-238).  Copies the topology, builds the edge list, derives
 * triangle records and samplers, and builds the LBVH.  Synchronises the stream
 * once (topology is read back to build edges). */
int psdr_scene_create(const psdr_scene_desc* desc, psdr_scene** out, void* stream);

/* Re-derive everything that depends on vertex positions, texels and emission
 * (triangle records, edge geometry, samplers, LBVH rebuild) from the buffers in
 * desc, which must have the same topology as at creation.  Fully asynchronous. */
int psdr_scene_update(psdr_scene* scene, const psdr_scene_desc* desc, void* stream);

int psdr_scene_destroy(psdr_scene* scene);

/* Orders the scene's outstanding work on its auxiliary streams (the boundary
 * prepass of a renderD forward whose backward has not run) before later work
 * on `stream`.  No reference counterpart: the reference runs everything on one
 * stream.  A graph replay that uses the scene calls it first. */
int psdr_scene_join(psdr_scene* scene, void* stream);

/* Makes `scene` use the per-render buffers (wavefront queues, rays, hits,
 * forward states, boundary scratch) and auxiliary streams of `with`; its own
 * are released (synchronising the device when no other scene shared them).
 * Scenes sharing them render one after another in stream order, as one
 * scene's calls do, and the buffers grow to the largest render: a batch of
 * scenes of uneven sizes rendered in turn holds one set instead of one per
 * scene.  Forward states stay per scene, four per workspace (least recently
 * used evicted; a backward whose state was evicted returns PSDR_ERR_STATE).
 * No reference counterpart: the reference renders one scene per call. */
int psdr_scene_share_workspace(psdr_scene* scene, psdr_scene* with, void* stream);

/* Counts: triangles, edges, LBVH depth (root = 0). */
int psdr_scene_info(const psdr_scene* scene, int32_t* n_tris, int32_t* n_edges, int32_t* bvh_depth);

/* ---- rendering ------------------------------------------------------------ */
/* PathIntegrator.renderC (src/path_integrator.cc:408-410 for one camera):
 * writes the (height, width, 3) fp32 image, row 0 = top.  With keep != 0 the
 * per-lane hit records are kept in the scene for psdr_render_backward
 * (the forward half of renderD, src/path_integrator.cc:412-414). */
int psdr_render_forward(psdr_scene* scene, const psdr_config* cfg, const psdr_camera* cam, float* out_rgb,
                        int32_t keep, void* stream);

/* The images of several cameras of one scene in one wavefront
 * (PathIntegrator::render_ renders the cameras of a scene one after another,
 * src/path_integrator.cc:385-403): every stage is one launch over the lanes of
 * all cameras, concatenated (camera c's lanes keep their own indices 0 ..
 * spp*H_c*W_c - 1, so each image equals psdr_render_forward's for that
 * camera).  out_rgb[c] receives camera c's (H_c, W_c, 3) image.  With keep != 0
 * each camera's forward state is kept for psdr_render_backward with that
 * camera.  At most PSDR_MAX_CAMERAS cameras per call (a scene keeps the forward
 * states of as many (config, camera) pairs). */
#define PSDR_MAX_CAMERAS 4
int psdr_render_forward_multi(psdr_scene* scene, const psdr_config* cfg, const psdr_camera* cams, int32_t n_cams,
                              float* const* out_rgb, int32_t keep, void* stream);

/* Terms of renderD's image derivative (src/path_integrator.cc:381-406). */
#define PSDR_TERM_INTERIOR 1 /* render_interior (path replay)                  */
#define PSDR_TERM_PRIMARY 2  /* render_primary_boundary (silhouette edges)     */
#define PSDR_TERM_DIRECT 4   /* render_direct_boundary (shadow edges)          */
#define PSDR_TERM_ALL 7

/* Backward of renderD (Enoki's FloatD.backward, src/python/torch_psdr.py:47-72)
 * for the image of a psdr_render_forward(keep=1) with the same camera and
 * config: accumulates dLoss/d(vertices), d(texels), d(light emit) and
 * d(camera parameters) (PSDR_CAMERA_GRAD floats) given dL_dI (height, width, 3).
 * The gradient buffers are ADDED to (zero them first).  Any gradient pointer
 * may be NULL.  `terms` selects PSDR_TERM_* (renderD uses PSDR_TERM_ALL).
 * Interior term by path replay; primary and direct-shadow boundary terms by
 * edge sampling with their hypercube preprocessing
 * (src/path_integrator.cc:188-353).  Camera parameters reach the image
 * derivative only through the ray origin: the camera vertex's direction to the
 * viewer (Scene::hit, src/scene.cc:134-136) and the primary boundary's origin
 * term (src/path_integrator.cc:240-244).  The scene keeps the forward state of
 * its last few distinct (camera, config) pairs, so the backward passes of a
 * multi-camera renderD find theirs without rendering again. */
int psdr_render_backward(psdr_scene* scene, const psdr_config* cfg, const psdr_camera* cam, const float* dL_dI,
                         float* grad_vertices, float* grad_texels, float* grad_emit, float* grad_camera,
                         int32_t terms, void* stream);

/* The same two entry points in the reference's launch shape, kept as the
 * measured baseline of the fused path (no reference counterpart): every depth
 * runs over all N lanes without compaction, each ray kind (light, BSDF,
 * extension sample) is traced by its own full-width launch followed by a host
 * synchronisation, as SceneOptix::hit synchronises after every launch
 * (src/scene_optix.cc:275), boundary samples are processed in lane order and the
 * gradient terms run one after another on the caller's stream.  Results equal
 * those of psdr_render_forward / psdr_render_backward. */
int psdr_render_forward_plain(psdr_scene* scene, const psdr_config* cfg, const psdr_camera* cam, float* out_rgb,
                              int32_t keep, void* stream);
int psdr_render_backward_plain(psdr_scene* scene, const psdr_config* cfg, const psdr_camera* cam,
                               const float* dL_dI, float* grad_vertices, float* grad_texels, float* grad_emit,
                               float* grad_camera, int32_t terms, void* stream);

/* ---- component entry points (tests, benchmarks) --------------------------- */
/* First n_draws PCG32 outputs of lanes [lane0, lane0+n_lanes) of Sampler
 * (src/sampler.h:11-17); out is device [n_lanes*n_draws]. */
int psdr_pcg_lane_draws(uint32_t lane0, int32_t n_lanes, int32_t n_draws, uint32_t* out, void* stream);

/* Camera rays of every lane of Sampler(spp*H*W) (src/camera.h:235-258); device
 * outputs o, d [N*3] and pixel index [N]. */
int psdr_camera_rays(const psdr_camera* cam, uint32_t spp, float* o, float* d, int32_t* pix, void* stream);

/* Host copies of the camera matrices (row-major 4x4): camera2world,
 * world2camera, camera2raster, raster2camera; and the camera position. */
int psdr_camera_matrices(const psdr_camera* cam, float* c2w, float* w2c, float* c2r, float* r2c, float* pos);

/* Host: the chain rule from camera-matrix adjoints to camera parameters that
 * psdr_render_backward applies on the device (Enoki's reverse mode through
 * look_at / perspective / orthographic and the matrix inverses,
 * src/camera.h:7-53, 108-124).  matrix_adjoints (25 floats): dLoss/d of
 * camera2world columns 0, 1, 3 (rows 0-2 each), world2camera rows 0, 1 and
 * raster2camera rows 0, 1; grad (PSDR_CAMERA_GRAD floats) is added to. */
int psdr_camera_param_vjp(const psdr_camera* cam, const float* matrix_adjoints, float* grad);

/* Closest hit (src/scene_optix.cc:240-277 + device_programs.cu:13-47) of n rays
 * against the scene: global triangle id (-1 on miss), t, barycentrics u, v
 * (weights of p1, p2; -1 on miss).  All device arrays; t may be NULL.  Grows
 * the scene's traversal-stack overflow buffer when the launch needs more. */
int psdr_scene_trace(psdr_scene* scene, const float* ro, const float* rd, int32_t n, int32_t* tri, float* t,
                     float* u, float* v, void* stream);

/* Largest ray count of psdr_scene_trace and psdr_scene_trace_noalloc: ray
 * buffers are indexed as 3 * i in 32-bit arithmetic. */
#define PSDR_NOALLOC_MAX_RAYS 715827882

/* psdr_scene_trace without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as psdr_scene_trace (bit-identical outputs); the
 * traversal stacks' overflow area is a fixed 32 MiB array of the library on
 * each device, used with a fixed launch of 1024 workgroups, instead of a scene
 * buffer grown on demand.  The first call on a device (n may be 0) looks the
 * array up, which may load the library's device code and with it the array;
 * make that call at set-up.  Calls no hipMalloc, hipFree or allocating library
 * routine.  n above
 * PSDR_NOALLOC_MAX_RAYS is refused with PSDR_ERR_ARG.  Every call on a device
 * shares that device's array: calls that overlap on one device (on different
 * streams or threads) race and must be ordered by the caller. */
int psdr_scene_trace_noalloc(psdr_scene* scene, const float* ro, const float* rd, int32_t n, int32_t* tri, float* t,
                             float* u, float* v, void* stream);

/* Derived data for parity tests (device outputs):
 * tri7 [n_tris*7] = normal(3), tangent(3), area;
 * edge_topo [n_edges*6] = v0, v1, left face, right face, left third vertex, right third vertex;
 * edge_geom [n_edges*16] = p0(3), p1(3), out(3), n_left(3), n_right(3), length. */
int psdr_scene_derived(const psdr_scene* scene, float* tri7, int32_t* edge_topo, float* edge_geom, void* stream);

/* Sampler tables (device outputs): light pick pmf/cmf over objects [n_objects],
 * edge pmf/cmf [n_edges], light power and area [n_lights]. */
int psdr_scene_samplers(const psdr_scene* scene, float* light_pmf, float* light_cmf, float* edge_pmf,
                        float* edge_cmf, float* light_power, float* light_area, void* stream);

/* LBVH of the scene (device outputs), as built by psdr_scene_create/update:
 * morton [n] by triangle, order [n] (sorted position -> triangle),
 * child [2(n-1)] (leaf k encoded as k | 0x80000000), parent [2n-1] (internal
 * nodes then leaves, root -1), box [(n-1)*12] (left lo/hi, right lo/hi). */
int psdr_scene_bvh(const psdr_scene* scene, uint32_t* morton, int32_t* order, int32_t* child, int32_t* parent,
                   float* box, void* stream);

/* Per-lane interior radiance of the last forward render (device [N*3]; the
 * lanes of every camera of a psdr_render_forward_multi, concatenated). */
int psdr_last_lane_radiance(const psdr_scene* scene, float* out, void* stream);

/* Counters of the last forward render: active paths entering each depth
 * (counts[0] = camera lanes, counts[d+1] = paths alive at depth d), and rays
 * traced; counts needs max_bounce + 2 entries.  Synchronises the stream. */
int psdr_last_render_stats(const psdr_scene* scene, int64_t* counts, int32_t n_counts, int64_t* rays_traced);

/* Live timing of the closest-hit traversal kernel (no reference counterpart;
 * used to measure it inside a timed region).  enable = 1 synchronises the
 * device, clears the totals and starts timing every traversal launch (at most
 * 4096 launches) two ways: a HIP event pair around the launch on its stream,
 * and, inside the kernel, the earliest workgroup start and latest workgroup end
 * on the device's constant-rate wall clock.  enable = 0 stops, synchronises and
 * returns the summed event durations (total_ms: they include time a launch
 * waits for CUs held by work on other streams), the summed in-kernel spans
 * (busy_ms: execution only, what a kernel trace reports as the dispatch's
 * duration), the number of launches and their algorithmic bytes (24 B ray read
 * + 12 B, or 16 B with t, hit written per ray).  Any output may be NULL.
 * Process-wide; not for concurrent callers. */
int psdr_trace_timing(int32_t enable, double* total_ms, int64_t* launches, int64_t* bytes, double* busy_ms);

/* Traversal counters (no reference counterpart; a profiling mode, also the
 * count of closest hits whose identity depends on rounding).  enable = 1
 * synchronises the device, clears the totals and makes every later traversal
 * launch use a counting instance of the kernel (same results, slower: it
 * prunes against the second-closest hit); enable = 0 stops, synchronises and
 * writes up to n_counts totals: [0] rays, [1] BVH nodes fetched, [2] triangles
 * tested (division-free prefilter), [3] exact ray-triangle tests, [4]
 * closest-hit rays whose closest and second-closest triangle distances lie
 * within 4 ulp (a near tie: another triangle test order or OptiX's arithmetic
 * could pick the other one), [5] closest-hit rays whose hit barycentrics lie
 * within 1e-6 of an edge (grazing: a neighbour could claim the hit), [6] rays
 * that hit, [7] closest-hit rays (visibility queries stop at the first
 * occluder and are not classified).  Process-wide; not for concurrent
 * callers. */
#define PSDR_TRACE_COUNTERS 8
int psdr_trace_counters(int32_t enable, int64_t* counts, int32_t n_counts);

/* Launch shapes of the hot kernels, chosen at run time by the autotuner
 * (torch_psdr_amd/tune.py) or by hand.  No counterpart in the reference, whose
 * launch shapes belong to OptiX and Enoki (src/scene_optix.cc:267,
 * src/path_integrator.cc:80-186).  Process-wide: a change applies to launches
 * issued afterwards (a captured graph keeps the shapes it was captured with)
 * and must not race with renders on other threads.  Results do not depend on
 * the shapes: images are bit-identical, gradients differ only by the order of
 * their float atomics. */
typedef struct psdr_tuning {
  int32_t trace_blocks_per_cu;   /* closest-hit kernel grid, blocks of 256 per CU, 1..64 (default 16) */
  int32_t trace_stack_lds;       /* traversal stack entries per lane kept in LDS: 8, 16 or 32 (default 16) */
  int32_t shade_blocks_per_cu;   /* wavefront shading kernel grid, 1..64 (default 8) */
  int32_t adjoint_blocks_per_cu; /* path-replay adjoint kernel grid, 1..64 (default 4) */
  int32_t direct_blocks_per_cu;  /* direct-boundary kernel grid, 1..64 (default 8) */
  int32_t tile_lanes;            /* lanes per wavefront tile of a forward render: 0 = all lanes at once
                                    (default), else >= PSDR_MIN_TILE_LANES; the queue, ray and hit buffers
                                    hold one tile, so their memory is O(tile) instead of O(spp*H*W) */
} psdr_tuning;
#define PSDR_MIN_TILE_LANES 65536

int psdr_get_tuning(psdr_tuning* out);
/* PSDR_ERR_ARG (nothing changed) when a field is outside its range. */
int psdr_set_tuning(const psdr_tuning* in);

#ifdef __cplusplus
}
#endif
#endif /* PSDR_HIP_H */
