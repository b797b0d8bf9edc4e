// Internal scene representation of the gfx950 path tracer (device data layout).
//
// HBM layout (all SoA / 16-byte records, global triangle order = object order):
//   tri_geom  [n_tris*5] float4 : {p0, area} {p1, -} {p2, -} {normal, -} {tangent, -}
//   tri_uv    [n_tris*6] float  : corner uvs (non-differentiable, src/mesh.cc:29-36)
//   leaf      [n_tris*3] float4 : LBVH-ordered {p0, tri id} {p1-p0, -} {p2-p0, -}
//   node      [(n-1)*3]  float4 : {lo_l.xyz, hi_l.x} {hi_l.yz, lo_r.xy} {lo_r.z, hi_r.xyz}, the
//                                 child ids in the first four planes' low bytes (traverse.h)
//   edge_geom [n_edges*16] float: p0 p1 out n_left n_right length (src/mesh.h:28-42)
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/psdr_hip.h"
#include "common.h"

namespace psdr {

// Error reporting: HIP failures become a PSDR_ERR_HIP status with a message.
struct Status {
  int code = PSDR_OK;
  std::string msg;
};
void set_error(int code, const std::string& msg);
#define PSDR_HIP_CHECK(expr)                                                              \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      ::psdr::set_error(PSDR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
      return PSDR_ERR_HIP;                                                                \
    }                                                                                     \
  } while (0)

// Stream capture (HIP graphs).  An entry point working on a stream that is
// being captured sets capture_active() for the duration of the call; buffers
// must not grow then (ensure() fails with PSDR_ERR_CAPTURE), and once any
// capture has happened in the process a buffer that grows retires its old
// allocation (kept until the process ends) instead of freeing it, since a graph
// may still point at it.
bool capture_active();
struct CaptureScope {
  bool prev;
  explicit CaptureScope(hipStream_t st);
  ~CaptureScope();
  CaptureScope(const CaptureScope&) = delete;
  CaptureScope& operator=(const CaptureScope&) = delete;
};

// status of a failed DevBuf::ensure (its message is already set)
inline int grow_failed() { return capture_active() ? PSDR_ERR_CAPTURE : PSDR_ERR_NOMEM; }

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int ensure(size_t n);
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Streams and events for running independent gradient terms side by side
// (created on first use, default priority).
struct AuxStreams {
  hipStream_t st[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  AuxStreams() = default;
  AuxStreams(const AuxStreams&) = delete;
  AuxStreams& operator=(const AuxStreams&) = delete;
  ~AuxStreams();
  int ensure();
};

// Object record as seen by kernels.
struct ObjectK {
  int mat, light;
  int alb_off, alb_h, alb_w;
  int rough_off, rough_h, rough_w;
  int tri_begin, tri_count;
  int grad_mask;
  int nrm_off, nrm_h, nrm_w;  // normal map (DiffuseBsdf), nrm_off < 0: none
};

// Light record as seen by kernels.
struct LightK {
  int obj;
  int tri_begin, tri_count;  // global triangle range of the light mesh
  int pmf_off;               // offset into light_tri_pmf/cmf
};

// Read-only view of a scene passed by value to render kernels.
// Most emitting triangles for which a BSDF-sampled ray tests the lights first
// (see SceneView::seek_leaf).
constexpr int SEEK_MAX = 16;

struct SceneView {
  const float4* tri_geom;
  const float* tri_uv;
  const int* tri_obj;
  const int* faces;
  const float4* leaf;
  const float4* node;
  int n_tris;
  int n_bvh;  // triangles in the LBVH: [0, n_bvh); the rest (an environment cube) are tested by every ray
  const ObjectK* objs;
  int n_objects;
  const LightK* lights;
  int n_lights;
  const float* light_emit;   // [n_lights*3]
  const float* light_power;  // [n_lights]
  const float* light_area;   // [n_lights]
  const float* lights_total; // [1]
  const float* light_tri_pmf;  // [n_light_tris]
  const float* light_tri_cmf;
  int n_light_tris;
  const float* obj_pmf;  // light pick over objects
  const float* obj_cmf;
  const float* texels;
  const float* vertices;
  const int* edge_topo;
  const float* edge_geom;
  const float* edge_pmf;
  const float* edge_cmf;
  int n_edges;
  // leaf records of every emitting triangle in the LBVH (n_seek of them), or
  // n_seek = -1: BSDF-sampled rays then resolve "which light, if any, is the
  // closest hit" by a full traversal (an environment light, whose cube every ray
  // hits, or more than SEEK_MAX emitting triangles)
  const float4* seek_leaf;
  int n_seek;
};

// Forward state of one renderD image kept for its backward pass: the
// interior's hit records and replay state, and the boundary terms' samples
// prepared during the forward render (with the scratch of their passes).  A
// scene keeps RENDER_SLOTS of them, one per distinct (config, camera), so the
// backward passes of a multi-camera renderD find their forward state.
struct RenderSlot {
  DevBuf hit_cache, replay;
  DevBuf b_rec, b_rays, b_hits, b_q0, b_LA, b_LB, b_rngA, b_rngB, b_state, b_tmp, b_counters, b_keys, b_vals, b_hist;
  DevBuf hc_mass, hc_pmf, hc_cmf;
  // primary-boundary wavefront scratch (separate from the interior render's)
  DevBuf p_qa, p_qb, p_rays, p_hits, p_rays2, p_hits2, p_spill;
  // direct-boundary scratch (separate from the primary pass's, so both can run at once)
  DevBuf d_state, d_mass, d_pmf, d_cmf, d_tmp, d_spill, d_keys, d_vals, d_hist, d_rec;
  DevBuf d_qidx, d_qval, d_qcnt;  // samples queued for the traversing kernel
  bool have_cache = false;  // hit_cache / replay hold the forward render of (cfg, cam)
  bool bprep = false;       // the boundary prepass for (cfg, cam) has been issued
  uint64_t N = 0;
  int K = 0, D = 0;
  psdr_config cfg{};
  psdr_camera cam{};
  const psdr_scene* owner = nullptr;  // the scene whose render this slot holds
  uint64_t used = 0;  // last-use stamp (least recently used slot is reused first)
};
constexpr int RENDER_SLOTS = 4;

// Per-render device buffers, forward states and auxiliary streams.  Every scene
// starts with its own; scenes may share one (psdr_scene_share_workspace), so
// that many scenes rendered one after another hold one set of buffers, sized
// for the largest render.  Calls on scenes sharing a workspace are ordered on
// the caller's stream as calls on one scene are.
struct Workspace {
  DevBuf lane_L, q_a, q_b, rays, hits, rays2, hits2, spill, counters_r, stats, cam_acc, jobs;
  // forward state per (scene, config, camera); cur = the slot the current call works on
  RenderSlot slots[RENDER_SLOTS];
  RenderSlot* cur = &slots[0];
  uint64_t clock = 0;
  AuxStreams aux;
  // work issued on the auxiliary streams and not yet joined into a caller's
  // stream; aux_captured: issued while the caller's stream was being captured
  bool aux_open = false, aux_captured = false;
  // last forward render (lane radiance and statistics entry points)
  const psdr_scene* last_owner = nullptr;
  uint64_t last_N = 0;  // lanes of every camera of the last forward render
  int last_K = 0, last_D = 0;
  int last_tiles = 1;   // lane tiles it ran (one block of counters each)
  psdr_config last_cfg{};
  psdr_camera last_cam{};
  int refs = 1;  // scenes using this workspace

  // slot holding the forward render of (owner, cfg, cam), or null
  RenderSlot* find_slot(const psdr_scene* owner, const psdr_config& cfg, const psdr_camera& cam);
  // slot for a new forward render: its own, a free one or the least recently used
  RenderSlot* claim_slot(const psdr_scene* owner, const psdr_config& cfg, const psdr_camera& cam);
  // drops the forward states of one scene
  void invalidate_slots(const psdr_scene* owner);
};

}  // namespace psdr

struct psdr_scene {
  int n_vertices = 0, n_tris = 0, n_objects = 0, n_lights = 0, n_texels = 0, n_edges = 0;
  int n_bvh = 0;  // triangles with edges and in the LBVH: all but an environment light's cube
  std::vector<psdr_object> objects;
  std::vector<psdr::ObjectK> objs_host;  // the bound object table
  std::vector<int> light_obj;
  // caller-owned device buffers of the last create/update
  const float* vertices = nullptr;
  const float* texels = nullptr;
  const float* light_emit = nullptr;
  // topology copies
  psdr::DevBuf faces, tri_uv, tri_obj, lights, edge_topo;
  // Device object tables, one per distinct content and never rewritten: a graph
  // captured with one table keeps reading it while eager calls bind another
  // (e.g. the same objects with other gradient masks).
  struct ObjTable {
    std::vector<psdr::ObjectK> host;
    psdr::DevBuf dev;
  };
  std::vector<std::unique_ptr<ObjTable>> obj_tables;
  const psdr::ObjectK* objs_dev = nullptr;
  // derived
  psdr::DevBuf tri_geom, edge_geom, edge_pmf, edge_cmf, light_tri_pmf, light_tri_cmf, light_power, light_area,
      lights_total, obj_pmf, obj_cmf, scan_tmp;
  // LBVH
  psdr::DevBuf morton, order, child, parent, box, leaf, node, counters, sort_keys, sort_vals, sort_hist, bounds,
      ranges, bvh_boxes;
  // per-render buffers, forward states, auxiliary streams (possibly shared)
  psdr::Workspace* ws = nullptr;
  // Reference launch shape, set for the duration of a *_plain call: every
  // depth's kernels run over all lanes (no compaction), each ray kind is traced
  // by its own launch followed by a host sync, boundary samples are processed in
  // lane order and the gradient terms run one after another.
  bool plain = false;
  int light_tri_total = 0;
  // emitting triangles' ranges (tri_begin, count) for the seek leaves, and their
  // total, or -1 (see SceneView::n_seek)
  std::vector<std::pair<int, int>> seek_ranges;
  int n_seek = -1;
  psdr::DevBuf seek_leaf;

  psdr_scene() = default;
  psdr_scene(const psdr_scene&) = delete;
  psdr_scene& operator=(const psdr_scene&) = delete;
  ~psdr_scene();
  psdr::SceneView view() const;
};

namespace psdr {
// scene.hip
int scene_derive(psdr_scene* s, hipStream_t st);
// Stable LSD radix sort of (key, value) pairs over the low 8*passes key bits
// (hist: radix_hist_words(n) words).  Buffers ping-pong per pass: the result is
// in (k0, v0) after an even number of passes, in (k1, v1) after an odd one.
size_t radix_hist_words(int n);
int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* hist, int n, int passes,
                     hipStream_t st);
int build_lbvh(psdr_scene* s, hipStream_t st);
// DiscreteSampler tables of n weights (single segment); tmp >= 2*(n/256+2) floats
int sampler_tables(const float* w, int n, float* pmf, float* cmf, float* total, float* tmp, hipStream_t st);
// camera (host, capi.cpp)
CameraK make_camera(const psdr_camera& c);
}  // namespace psdr
