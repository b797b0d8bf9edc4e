// C ABI of the gfx950 path tracer (include/psdr_hip.h).  Host-side scene
// lifecycle, camera matrices and argument checking; no exception crosses the
// boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "cam_grad.h"
#include "render.h"
#include "scene.h"
#include "traverse.h"

namespace psdr {

thread_local std::string g_last_error;

void set_error(int code, const std::string& msg) {
  (void)code;
  g_last_error = msg;
}

namespace {
thread_local bool t_capture = false;
std::atomic<bool> g_graph_seen{false};
std::mutex g_retired_mu;
std::vector<void*> g_retired;  // allocations a captured graph may still use
}  // namespace

bool capture_active() { return t_capture; }

CaptureScope::CaptureScope(hipStream_t st) : prev(t_capture) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    (void)hipGetLastError();
    cs = hipStreamCaptureStatusNone;
  }
  t_capture = cs == hipStreamCaptureStatusActive;
  if (t_capture) g_graph_seen = true;
}
CaptureScope::~CaptureScope() { t_capture = prev; }

int DevBuf::ensure(size_t n) {
  if (n <= bytes && p) return PSDR_OK;
  if (t_capture) {
    set_error(PSDR_ERR_CAPTURE, "a scene buffer would grow to " + std::to_string(n) +
                                    " bytes during graph capture: run the captured calls once eagerly, with the "
                                    "same sizes, before capturing them");
    return PSDR_ERR_CAPTURE;
  }
  if (p && g_graph_seen) {
    std::lock_guard<std::mutex> lk(g_retired_mu);
    g_retired.push_back(p);
    p = nullptr;
    bytes = 0;
  }
  release();
  if (n == 0) n = 16;
  if (hipMalloc(&p, n) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    set_error(PSDR_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed");
    return PSDR_ERR_NOMEM;
  }
  bytes = n;
  return PSDR_OK;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

AuxStreams::~AuxStreams() {
  for (int i = 0; i < 2; ++i) {
    if (st[i]) (void)hipStreamDestroy(st[i]);
    if (join[i]) (void)hipEventDestroy(join[i]);
  }
  if (fork) (void)hipEventDestroy(fork);
}

int AuxStreams::ensure() {
  if (fork) return PSDR_OK;
  if (t_capture) {
    set_error(PSDR_ERR_CAPTURE, "auxiliary streams are created on first use: run the captured calls once eagerly "
                                "before capturing them");
    return PSDR_ERR_CAPTURE;
  }
  for (int i = 0; i < 2; ++i) {
    PSDR_HIP_CHECK(hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking));
    PSDR_HIP_CHECK(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
  }
  PSDR_HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  return PSDR_OK;
}

// ---------------------------------------------------------------------------
// Camera matrices (src/camera.h:7-17, 33-53, 108-124; near/far src/psdr.cc:27-28)
// in fp32; row-major m[r][c]; products summed k = 0..3 left to right.
// ---------------------------------------------------------------------------
static Mat4 mat_mul(const Mat4& a, const Mat4& b) {
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      float s = a.m[i][0] * b.m[0][j];
      s = s + a.m[i][1] * b.m[1][j];
      s = s + a.m[i][2] * b.m[2][j];
      s = s + a.m[i][3] * b.m[3][j];
      r.m[i][j] = s;
    }
  return r;
}
// Inverse by the 2x2-minor (Laplace) expansion over row pairs (0,1) and (2,3).
static Mat4 mat_inverse(const Mat4& M) {
  const float(*a)[4] = M.m;
  const float s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
  const float s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
  const float s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
  const float s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
  const float s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
  const float s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
  const float c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
  const float c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
  const float c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
  const float c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
  const float c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
  const float c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
  const float det = ((((s0 * c5 - s1 * c4) + s2 * c3) + s3 * c2) - s4 * c1) + s5 * c0;
  const float id = 1.0f / det;
  Mat4 b;
  b.m[0][0] = ((a[1][1] * c5 - a[1][2] * c4) + a[1][3] * c3) * id;
  b.m[0][1] = ((-a[0][1] * c5 + a[0][2] * c4) - a[0][3] * c3) * id;
  b.m[0][2] = ((a[3][1] * s5 - a[3][2] * s4) + a[3][3] * s3) * id;
  b.m[0][3] = ((-a[2][1] * s5 + a[2][2] * s4) - a[2][3] * s3) * id;
  b.m[1][0] = ((-a[1][0] * c5 + a[1][2] * c2) - a[1][3] * c1) * id;
  b.m[1][1] = ((a[0][0] * c5 - a[0][2] * c2) + a[0][3] * c1) * id;
  b.m[1][2] = ((-a[3][0] * s5 + a[3][2] * s2) - a[3][3] * s1) * id;
  b.m[1][3] = ((a[2][0] * s5 - a[2][2] * s2) + a[2][3] * s1) * id;
  b.m[2][0] = ((a[1][0] * c4 - a[1][1] * c2) + a[1][3] * c0) * id;
  b.m[2][1] = ((-a[0][0] * c4 + a[0][1] * c2) - a[0][3] * c0) * id;
  b.m[2][2] = ((a[3][0] * s4 - a[3][1] * s2) + a[3][3] * s0) * id;
  b.m[2][3] = ((-a[2][0] * s4 + a[2][1] * s2) - a[2][3] * s0) * id;
  b.m[3][0] = ((-a[1][0] * c3 + a[1][1] * c1) - a[1][2] * c0) * id;
  b.m[3][1] = ((a[0][0] * c3 - a[0][1] * c1) + a[0][2] * c0) * id;
  b.m[3][2] = ((-a[3][0] * s3 + a[3][1] * s1) - a[3][2] * s0) * id;
  b.m[3][3] = ((a[2][0] * s3 - a[2][1] * s1) + a[2][2] * s0) * id;
  return b;
}

CameraK make_camera(const psdr_camera& c) {
  CameraK k;
  k.H = c.height;
  k.W = c.width;
  const f3 o{c.look_from[0], c.look_from[1], c.look_from[2]};
  const f3 t{c.look_at[0], c.look_at[1], c.look_at[2]};
  const f3 up{c.up[0], c.up[1], c.up[2]};
  const f3 dir = normalize(o - t);
  const f3 right = cross(normalize(up), dir);
  const f3 nup = cross(dir, right);
  const f3 cols[4] = {right, nup, dir, o};
  for (int r = 0; r < 4; ++r)
    for (int col = 0; col < 4; ++col) k.c2w.m[r][col] = r < 3 ? get(cols[col], r) : (col == 3 ? 1.f : 0.f);
  k.w2c = mat_inverse(k.c2w);
  const float n = -0.01f, f = -1000.f;
  Mat4 P;
  std::memset(&P, 0, sizeof(P));
  k.ortho = c.kind == PSDR_CAMERA_ORTHOGRAPHIC ? 1 : 0;
  if (k.ortho) {
    // ortho_mat (src/camera.h:302-312) = scale * translate of orthographic()
    // (src/camera.h:19-31): half extents view_width / 2 and H * (view_width / 2) / W
    const float hw = c.view_width / 2.0f;
    const float hh = (static_cast<float>(c.height) * hw) / static_cast<float>(c.width);
    const float l = -hw, r = hw, b = -hh, t = hh;
    Mat4 Sc, Tr;
    std::memset(&Sc, 0, sizeof(Sc));
    std::memset(&Tr, 0, sizeof(Tr));
    for (int i = 0; i < 4; ++i) Tr.m[i][i] = 1.f;
    Tr.m[0][3] = -(l + r) / 2.0f;
    Tr.m[1][3] = -(b + t) / 2.0f;
    Tr.m[2][3] = -(n + f) / 2.0f;
    Sc.m[0][0] = 2.0f / (r - l);
    Sc.m[1][1] = 2.0f / (t - b);
    Sc.m[2][2] = 2.0f / (n - f);
    Sc.m[3][3] = 1.f;
    P = mat_mul(Sc, Tr);
  } else {
    const float inv_nf = 1.0f / (n - f);
    float sh, ch;
    psdr_sincos(c.vfov * 0.5f, &sh, &ch);
    const float cot = ch / sh;
    P.m[0][0] = -cot;
    P.m[1][1] = -cot;
    P.m[2][2] = (n + f) * inv_nf;
    P.m[2][3] = ((-2.f * f) * n) * inv_nf;
    P.m[3][2] = 1.f;
  }
  Mat4 S, T;
  std::memset(&S, 0, sizeof(S));
  std::memset(&T, 0, sizeof(T));
  S.m[0][0] = 0.5f * static_cast<float>(c.width);
  S.m[1][1] = -0.5f * static_cast<float>(c.height);
  S.m[2][2] = 1.f;
  S.m[3][3] = 1.f;
  for (int i = 0; i < 4; ++i) T.m[i][i] = 1.f;
  T.m[0][3] = 1.f;
  T.m[1][3] = -1.f;
  k.c2r = mat_mul(mat_mul(S, T), P);
  k.r2c = mat_inverse(k.c2r);
  k.pos = xform_pos(k.c2w, f3{0.f, 0.f, 0.f});
  k.look = xform_dir(k.c2w, f3{0.f, 0.f, -1.f});
  return k;
}

// Canonical edge list (src/mesh.cc:97-204 with (i<j)-sorted order): half-edges
// keyed by (min, max) vertex id, stable-sorted; the first face is the left one,
// the last face sharing the key is the right one; an unshared edge reuses the
// left face.
static std::vector<int> edge_topology(const std::vector<int>& faces, int nf) {
  struct HalfEdge {
    uint64_t key;
    int face, k;
  };
  std::vector<HalfEdge> he;
  he.reserve(static_cast<size_t>(nf) * 3);
  for (int f = 0; f < nf; ++f) {
    const int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
    const int trip[3][3] = {{a, b, c}, {b, c, a}, {c, a, b}};
    for (int s = 0; s < 3; ++s) {
      int i = trip[s][0], j = trip[s][1];
      if (j < i) std::swap(i, j);
      he.push_back({(static_cast<uint64_t>(static_cast<uint32_t>(i)) << 32) | static_cast<uint32_t>(j), f,
                    trip[s][2]});
    }
  }
  std::stable_sort(he.begin(), he.end(), [](const HalfEdge& x, const HalfEdge& y) { return x.key < y.key; });
  std::vector<int> out;
  out.reserve(he.size() * 3);
  for (size_t s = 0; s < he.size();) {
    size_t e = s;
    while (e + 1 < he.size() && he[e + 1].key == he[s].key) ++e;
    out.push_back(static_cast<int>(he[s].key >> 32));
    out.push_back(static_cast<int>(he[s].key & 0xffffffffu));
    out.push_back(he[s].face);
    out.push_back(e > s ? he[e].face : he[s].face);
    out.push_back(he[s].k);
    out.push_back(e > s ? he[e].k : he[s].k);
    s = e + 1;
  }
  return out;
}

static int check_desc(const psdr_scene_desc* d) {
  if (!d) {
    set_error(PSDR_ERR_ARG, "scene desc is NULL");
    return PSDR_ERR_ARG;
  }
  if (d->n_vertices < 0 || d->n_tris < 0 || d->n_objects < 0 || d->n_lights < 0 || d->n_texels < 0) {
    set_error(PSDR_ERR_ARG, "negative count in scene desc");
    return PSDR_ERR_ARG;
  }
  if (d->n_tris > PSDR_MAX_TRIS) {
    set_error(PSDR_ERR_ARG, "scene desc: more than PSDR_MAX_TRIS triangles");
    return PSDR_ERR_ARG;
  }
  if (d->n_tris > 0 && (!d->vertices || !d->faces || !d->tri_uv || !d->tri_obj)) {
    set_error(PSDR_ERR_ARG, "scene desc: missing geometry buffer");
    return PSDR_ERR_ARG;
  }
  if (d->n_objects > 0 && !d->objects) {
    set_error(PSDR_ERR_ARG, "scene desc: objects is NULL");
    return PSDR_ERR_ARG;
  }
  if (d->n_lights > 0 && (!d->light_emit || !d->light_obj)) {
    set_error(PSDR_ERR_ARG, "scene desc: light arrays are NULL");
    return PSDR_ERR_ARG;
  }
  for (int o = 0; o < d->n_objects; ++o) {
    const psdr_object& ob = d->objects[o];
    if (ob.tri_begin < 0 || ob.tri_count < 0 || ob.tri_begin + ob.tri_count > d->n_tris) {
      set_error(PSDR_ERR_ARG, "object " + std::to_string(o) + ": triangle range out of bounds");
      return PSDR_ERR_ARG;
    }
    if (ob.light >= d->n_lights || ob.material < 0 || ob.material > PSDR_MAT_ENV_SH9) {
      set_error(PSDR_ERR_ARG, "object " + std::to_string(o) + ": bad light index or material");
      return PSDR_ERR_ARG;
    }
    const bool env = ob.material == PSDR_MAT_ENV_MAP || ob.material == PSDR_MAT_ENV_SH9;
    if (env && (o != d->n_objects - 1 || ob.light < 0 || ob.tri_count != 12 ||
                ob.tri_begin + ob.tri_count != d->n_tris)) {
      set_error(PSDR_ERR_ARG, "object " + std::to_string(o) +
                                  ": an environment light's cube must be the last object, a light, 12 triangles");
      return PSDR_ERR_ARG;
    }
    const int64_t alb = ob.material == PSDR_MAT_ENV_SH9 ? 27 : (ob.albedo_h == 0 ? 3 : 3ll * ob.albedo_h * ob.albedo_w);
    if (ob.material != PSDR_MAT_LIGHT && (ob.albedo_offset < 0 || ob.albedo_offset + alb > d->n_texels)) {
      set_error(PSDR_ERR_ARG, "object " + std::to_string(o) + ": albedo texture out of bounds");
      return PSDR_ERR_ARG;
    }
    const int64_t rough = ob.rough_h == 0 ? 1 : 1ll * ob.rough_h * ob.rough_w;
    if (ob.material == PSDR_MAT_DIFFUSE_BSDF && (ob.rough_offset < 0 || ob.rough_offset + rough > d->n_texels)) {
      set_error(PSDR_ERR_ARG, "object " + std::to_string(o) + ": roughness texture out of bounds");
      return PSDR_ERR_ARG;
    }
    const int64_t nrm = ob.normal_h == 0 ? 3 : 3ll * ob.normal_h * ob.normal_w;
    if (ob.material == PSDR_MAT_DIFFUSE_BSDF && ob.normal_offset != -1 &&
        (ob.normal_offset < 0 || ob.normal_h < 0 || ob.normal_w < 0 || ob.normal_offset + nrm > d->n_texels)) {
      set_error(PSDR_ERR_ARG, "object " + std::to_string(o) + ": normal map out of bounds");
      return PSDR_ERR_ARG;
    }
  }
  for (int l = 0; l < d->n_lights; ++l) {
    const int o = d->light_obj[l];
    if (o < 0 || o >= d->n_objects || d->objects[o].light != l || d->objects[o].tri_count <= 0) {
      set_error(PSDR_ERR_ARG, "light " + std::to_string(l) + ": bad object");
      return PSDR_ERR_ARG;
    }
  }
  return PSDR_OK;
}

}  // namespace psdr

using namespace psdr;

static bool same_render(const RenderSlot& r, const psdr_scene* owner, const psdr_config& cfg,
                        const psdr_camera& cam) {
  return r.owner == owner && std::memcmp(&r.cfg, &cfg, sizeof(cfg)) == 0 &&
         std::memcmp(&r.cam, &cam, sizeof(cam)) == 0;
}

RenderSlot* Workspace::find_slot(const psdr_scene* owner, const psdr_config& cfg, const psdr_camera& cam) {
  for (RenderSlot& r : slots)
    if (r.have_cache && same_render(r, owner, cfg, cam)) {
      r.used = ++clock;
      return &r;
    }
  return nullptr;
}

RenderSlot* Workspace::claim_slot(const psdr_scene* owner, const psdr_config& cfg, const psdr_camera& cam) {
  // the slot last used for this render, valid or not (its buffers already have
  // the right sizes: a training loop's scene update must not move renders to a
  // new slot), else one never used, else the least recently used
  RenderSlot* best = nullptr;
  for (RenderSlot& r : slots)
    if (r.used && same_render(r, owner, cfg, cam)) {
      best = &r;
      break;
    }
  if (!best)
    for (RenderSlot& r : slots)
      if (!best || r.used < best->used) best = &r;
  best->used = ++clock;
  best->cfg = cfg;
  best->cam = cam;
  best->owner = owner;
  return best;
}

void Workspace::invalidate_slots(const psdr_scene* owner) {
  for (RenderSlot& r : slots)
    if (r.owner == owner) {
      r.have_cache = false;
      r.bprep = false;
    }
}

psdr_scene::~psdr_scene() {
  if (!ws) return;
  if (--ws->refs == 0) {
    delete ws;
    return;
  }
  // the shared workspace forgets this scene's renders
  ws->invalidate_slots(this);
  for (RenderSlot& r : ws->slots)
    if (r.owner == this) r.owner = nullptr;
  if (ws->last_owner == this) ws->last_owner = nullptr;
}

SceneView psdr_scene::view() const {
  SceneView v;
  v.tri_geom = tri_geom.as<float4>();
  v.tri_uv = tri_uv.as<float>();
  v.tri_obj = tri_obj.as<int>();
  v.faces = faces.as<int>();
  v.leaf = leaf.as<float4>();
  v.node = node.as<float4>();
  v.n_tris = n_tris;
  v.n_bvh = n_bvh;
  v.objs = objs_dev;
  v.n_objects = n_objects;
  v.lights = lights.as<LightK>();
  v.n_lights = n_lights;
  v.light_emit = light_emit;
  v.light_power = light_power.as<float>();
  v.light_area = light_area.as<float>();
  v.lights_total = lights_total.as<float>();
  v.light_tri_pmf = light_tri_pmf.as<float>();
  v.light_tri_cmf = light_tri_cmf.as<float>();
  v.n_light_tris = light_tri_total;
  v.obj_pmf = obj_pmf.as<float>();
  v.obj_cmf = obj_cmf.as<float>();
  v.texels = texels;
  v.vertices = vertices;
  v.edge_topo = edge_topo.as<int>();
  v.edge_geom = edge_geom.as<float>();
  v.edge_pmf = edge_pmf.as<float>();
  v.edge_cmf = edge_cmf.as<float>();
  v.n_edges = n_edges;
  v.seek_leaf = seek_leaf.as<float4>();
  v.n_seek = n_seek;
  return v;
}

#define PSDR_TRY(expr)        \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != PSDR_OK) return rc_; \
  } while (0)

static hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

static int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st) {
  PSDR_TRY(b.ensure(bytes));
  if (bytes) PSDR_HIP_CHECK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
  return PSDR_OK;
}

// Points the scene at the device copy of objs_host, uploading it once per
// distinct content.
static int bind_object_table(psdr_scene* s, hipStream_t st) {
  const size_t bytes = sizeof(ObjectK) * s->objs_host.size();
  for (const auto& t : s->obj_tables)
    if (t->host.size() == s->objs_host.size() && std::memcmp(t->host.data(), s->objs_host.data(), bytes) == 0) {
      s->objs_dev = t->dev.as<ObjectK>();
      return PSDR_OK;
    }
  std::unique_ptr<psdr_scene::ObjTable> t(new (std::nothrow) psdr_scene::ObjTable());
  if (!t) {
    set_error(PSDR_ERR_NOMEM, "host allocation failed");
    return PSDR_ERR_NOMEM;
  }
  t->host = s->objs_host;
  PSDR_TRY(upload(t->dev, t->host.data(), bytes, st));
  s->objs_dev = t->dev.as<ObjectK>();
  s->obj_tables.push_back(std::move(t));
  return PSDR_OK;
}

static ObjectK object_k(const psdr_object& b) {
  return ObjectK{b.material,  b.light,   b.albedo_offset, b.albedo_h,  b.albedo_w,
                 b.rough_offset, b.rough_h, b.rough_w,     b.tri_begin, b.tri_count,
                 b.grad_mask, b.material == PSDR_MAT_DIFFUSE_BSDF ? b.normal_offset : -1,
                 b.normal_h,  b.normal_w};
}

// Points the scene at the caller's vertex / texel / emission buffers.  On an
// update the object table may change in everything but its triangle ranges and
// light slots (texture offsets and sizes, material, which parameters take
// gradients); a changed table is uploaded again.
static int scene_set_buffers(psdr_scene* s, const psdr_scene_desc* d, hipStream_t st) {
  if (d->n_tris != s->n_tris || d->n_vertices != s->n_vertices || d->n_objects != s->n_objects ||
      d->n_lights != s->n_lights) {
    set_error(PSDR_ERR_ARG, "scene update: counts differ from the scene at creation");
    return PSDR_ERR_ARG;
  }
  bool changed = false;
  for (int o = 0; o < d->n_objects; ++o) {
    const psdr_object& a = s->objects[o];
    const psdr_object& b = d->objects[o];
    if (a.tri_begin != b.tri_begin || a.tri_count != b.tri_count || a.light != b.light) {
      set_error(PSDR_ERR_ARG, "scene update: object " + std::to_string(o) + " changed its triangles or light slot");
      return PSDR_ERR_ARG;
    }
    changed = changed || std::memcmp(&a, &b, sizeof(psdr_object)) != 0;
  }
  if (changed) {
    s->objects.assign(d->objects, d->objects + d->n_objects);
    s->objs_host.resize(s->objects.size());
    for (size_t o = 0; o < s->objects.size(); ++o) s->objs_host[o] = object_k(s->objects[o]);
    PSDR_TRY(bind_object_table(s, st));
  }
  s->vertices = d->vertices;
  s->texels = d->texels;
  s->light_emit = d->light_emit;
  s->n_texels = d->n_texels;
  return PSDR_OK;
}

static int scene_create_impl(const psdr_scene_desc* d, psdr_scene* s, hipStream_t st) {
  PSDR_TRY(check_desc(d));
  s->n_vertices = d->n_vertices;
  s->n_tris = d->n_tris;
  s->n_objects = d->n_objects;
  s->n_lights = d->n_lights;
  s->n_texels = d->n_texels;
  s->objects.assign(d->objects, d->objects + d->n_objects);
  s->light_obj.assign(d->light_obj, d->light_obj + d->n_lights);
  const int nt = d->n_tris;
  // topology copies
  PSDR_TRY(s->faces.ensure(12ull * std::max(nt, 1)));
  PSDR_TRY(s->tri_uv.ensure(24ull * std::max(nt, 1)));
  PSDR_TRY(s->tri_obj.ensure(4ull * std::max(nt, 1)));
  if (nt) {
    PSDR_HIP_CHECK(hipMemcpyAsync(s->faces.p, d->faces, 12ull * nt, hipMemcpyDeviceToDevice, st));
    PSDR_HIP_CHECK(hipMemcpyAsync(s->tri_uv.p, d->tri_uv, 24ull * nt, hipMemcpyDeviceToDevice, st));
    PSDR_HIP_CHECK(hipMemcpyAsync(s->tri_obj.p, d->tri_obj, 4ull * nt, hipMemcpyDeviceToDevice, st));
  }
  std::vector<int> hfaces(3ull * nt), hobj(nt);
  if (nt) {
    PSDR_HIP_CHECK(hipMemcpyAsync(hfaces.data(), d->faces, 12ull * nt, hipMemcpyDeviceToHost, st));
    PSDR_HIP_CHECK(hipMemcpyAsync(hobj.data(), d->tri_obj, 4ull * nt, hipMemcpyDeviceToHost, st));
  }
  PSDR_HIP_CHECK(hipStreamSynchronize(st));
  for (int f = 0; f < nt; ++f)
    for (int k = 0; k < 3; ++k)
      if (hfaces[3 * f + k] < 0 || hfaces[3 * f + k] >= d->n_vertices) {
        set_error(PSDR_ERR_ARG, "face " + std::to_string(f) + " references a vertex out of range");
        return PSDR_ERR_ARG;
      }
  // kernels index the object table with tri_obj[t] and assume t lies in that
  // object's triangle range
  for (int t = 0; t < nt; ++t) {
    const int o = hobj[t];
    if (o < 0 || o >= d->n_objects || t < d->objects[o].tri_begin ||
        t >= d->objects[o].tri_begin + d->objects[o].tri_count) {
      set_error(PSDR_ERR_ARG, "triangle " + std::to_string(t) + ": tri_obj " + std::to_string(o) +
                                  " is not an object whose triangle range holds it");
      return PSDR_ERR_ARG;
    }
  }
  // an environment light's cube (the last object) has no edges and stays out of the LBVH
  s->n_bvh = nt;
  if (d->n_objects > 0) {
    const psdr_object& last = d->objects[d->n_objects - 1];
    if (last.material == PSDR_MAT_ENV_MAP || last.material == PSDR_MAT_ENV_SH9) s->n_bvh = last.tri_begin;
  }
  std::vector<int> topo = edge_topology(hfaces, s->n_bvh);
  s->n_edges = static_cast<int>(topo.size() / 6);
  // objects / lights
  s->objs_host.resize(d->n_objects);
  for (int o = 0; o < d->n_objects; ++o) s->objs_host[o] = object_k(d->objects[o]);
  // the emitting triangles a BSDF-sampled ray tests before it traverses
  s->seek_ranges.clear();
  s->n_seek = -1;
  if (s->n_bvh == nt) {
    int cnt = 0;
    for (int o = 0; o < d->n_objects; ++o)
      if (d->objects[o].light >= 0 && d->objects[o].tri_count > 0) {
        s->seek_ranges.emplace_back(d->objects[o].tri_begin, d->objects[o].tri_count);
        cnt += d->objects[o].tri_count;
      }
    if (cnt <= SEEK_MAX) s->n_seek = cnt;
  }
  PSDR_TRY(s->seek_leaf.ensure(48ull * SEEK_MAX));
  std::vector<LightK> lk(d->n_lights);
  int off = 0;
  for (int l = 0; l < d->n_lights; ++l) {
    const psdr_object& b = d->objects[d->light_obj[l]];
    lk[l] = LightK{d->light_obj[l], b.tri_begin, b.tri_count, off};
    off += b.tri_count;
  }
  s->light_tri_total = off;
  const int ne = s->n_edges, nl = d->n_lights, no = d->n_objects;
  // scan scratch: ints [seg_begin nl, seg_count nl, n_objects, n_edges, pad] + floats
  std::vector<int> hdr(2 * nl + 4, 0);
  hdr[2 * nl] = no;
  hdr[2 * nl + 1] = ne;
  const size_t wmax = std::max<size_t>({static_cast<size_t>(off), static_cast<size_t>(ne), static_cast<size_t>(no),
                                        1});
  const size_t tmp_floats = 2 * (wmax / 256 + 2 + static_cast<size_t>(nl) + 2) + 16;
  PSDR_TRY(s->scan_tmp.ensure(4 * (hdr.size() + wmax + tmp_floats)));
  PSDR_HIP_CHECK(hipMemcpyAsync(s->scan_tmp.p, hdr.data(), 4 * hdr.size(), hipMemcpyHostToDevice, st));
  PSDR_TRY(bind_object_table(s, st));
  PSDR_TRY(upload(s->lights, lk.data(), sizeof(LightK) * lk.size(), st));
  PSDR_TRY(upload(s->edge_topo, topo.data(), 4 * topo.size(), st));
  // derived buffers
  const size_t n1 = static_cast<size_t>(std::max(nt, 1));
  PSDR_TRY(s->tri_geom.ensure(80 * n1));
  PSDR_TRY(s->edge_geom.ensure(64ull * std::max(ne, 1)));
  PSDR_TRY(s->edge_pmf.ensure(4ull * std::max(ne, 1)));
  PSDR_TRY(s->edge_cmf.ensure(4ull * std::max(ne, 1)));
  PSDR_TRY(s->light_tri_pmf.ensure(4ull * std::max(off, 1)));
  PSDR_TRY(s->light_tri_cmf.ensure(4ull * std::max(off, 1)));
  PSDR_TRY(s->light_power.ensure(4ull * std::max(nl, 1)));
  PSDR_TRY(s->light_area.ensure(4ull * std::max(nl, 1)));
  PSDR_TRY(s->lights_total.ensure(16));
  PSDR_TRY(s->obj_pmf.ensure(4ull * std::max(no, 1)));
  PSDR_TRY(s->obj_cmf.ensure(4ull * std::max(no, 1)));
  PSDR_TRY(s->morton.ensure(4 * n1));
  PSDR_TRY(s->order.ensure(4 * n1));
  PSDR_TRY(s->child.ensure(8 * n1));
  PSDR_TRY(s->parent.ensure(8 * n1));
  PSDR_TRY(s->box.ensure(48 * n1));
  PSDR_TRY(s->leaf.ensure(48 * n1));
  PSDR_TRY(s->node.ensure(48 * n1));
  PSDR_TRY(s->counters.ensure(4 * n1));
  PSDR_TRY(s->sort_keys.ensure(8 * n1));
  PSDR_TRY(s->sort_vals.ensure(8 * n1));
  PSDR_TRY(s->sort_hist.ensure(4ull * radix_hist_words(n1)));
  PSDR_TRY(s->ranges.ensure(8ull * n1));
  // leaf boxes [n], internal-node boxes [n-1], 64-leaf chunk boxes, 6 floats each
  PSDR_TRY(s->bvh_boxes.ensure(24ull * (2ull * n1 + n1 / 64 + 1)));
  PSDR_TRY(s->bounds.ensure(4ull * (8 + 6 * 1024)));
  PSDR_TRY(scene_set_buffers(s, d, st));
  PSDR_TRY(scene_derive(s, st));
  PSDR_HIP_CHECK(hipStreamSynchronize(st));
  return PSDR_OK;
}

static bool cfg_ok(const psdr_config* c) {
  if (!c) return false;
  return c->n_pass > 0 && c->spp_interior > 0 && c->spp_primary_edge > 0 && c->spp_secondary_edge > 0 &&
         c->max_bounce > 0 && c->max_bounce <= 64 && c->primary_edge_preprocess_rounds > 0 &&
         c->secondary_edge_preprocess_rounds > 0 && c->primary_edge_hypercube_resolution > 0 &&
         c->secondary_edge_hypercube_resolution[0] > 0 && c->secondary_edge_hypercube_resolution[1] > 0 &&
         c->secondary_edge_hypercube_resolution[2] > 0;
}
static bool cam_ok(const psdr_camera* c) {
  if (!c || c->height <= 0 || c->width <= 0) return false;
  if (c->kind == PSDR_CAMERA_PERSPECTIVE) return true;
  return c->kind == PSDR_CAMERA_ORTHOGRAPHIC && c->view_width > 0.f;
}

// Size limits of one render.  The wavefront kernels index ray buffers in 32-bit
// arithmetic as 3 * i with i < lanes * (rays per lane): the interior render has
// N = spp * H * W lanes with L + B + 1 rays each, the primary boundary
// max(spp_primary_edge * H * W, hypercube cells) lanes with L + 2 rays each (one
// light, one BSDF and one extension ray per depth, two camera-side rays), the
// direct boundary spp_secondary_edge * H * W lanes over at most 2^28 cells.
// Checked before any buffer is touched, so it needs no device.
static int check_render_sizes(const psdr_config& c, const psdr_camera& cam) {
  constexpr int64_t lim = 0x7fffffff;
  const int64_t px = static_cast<int64_t>(cam.height) * cam.width;
  const char* why = nullptr;
  if (c.mis_light_samples > 4096 || c.mis_bsdf_samples > 4096) {
    why = "mis_light_samples and mis_bsdf_samples must be at most 4096";
  } else {
    const int64_t ki = 1 + int64_t{c.mis_light_samples} + c.mis_bsdf_samples;
    const int64_t kp = 2 + int64_t{c.mis_light_samples};
    const int64_t ni = int64_t{c.spp_interior} * px, np = int64_t{c.spp_primary_edge} * px;
    const int64_t nd = int64_t{c.spp_secondary_edge} * px;
    const int64_t cells = int64_t{c.secondary_edge_hypercube_resolution[0]} *
                          c.secondary_edge_hypercube_resolution[1] * c.secondary_edge_hypercube_resolution[2];
    if (ni > lim / 4 || 3 * ki * ni > lim)
      why = "interior: 3 * spp_interior * H * W * (mis_light_samples + mis_bsdf_samples + 1) exceeds 2^31 - 1";
    else if (3 * kp * std::max<int64_t>(np, c.primary_edge_hypercube_resolution) > lim)
      why = "primary boundary: 3 * max(spp_primary_edge * H * W, primary_edge_hypercube_resolution) * "
            "(mis_light_samples + 2) exceeds 2^31 - 1";
    else if (nd > lim / 8 || cells > lim / 8)
      why = "direct boundary: spp_secondary_edge * H * W and the hypercube cells must each be below 2^28";
  }
  if (!why) return PSDR_OK;
  set_error(PSDR_ERR_ARG, std::string("render: sizes out of range: ") + why);
  return PSDR_ERR_ARG;
}

// Scratch of psdr_scene_trace_noalloc: the overflow area of the traversal
// stacks (entries 32..63 of each lane's stack; 0..31 live in LDS) of
// NOALLOC_BLOCKS persistent workgroups, 32 MiB per device, held by every device
// the library runs on until the process ends.  Nothing in it must start at
// zero: a lane writes a slot (push) before it reads it (pop) in every
// traversal.  Host code uses the per-device address from hipGetSymbolAddress,
// never the array's name (that is a host-side stand-in, not a device address).
namespace {
constexpr int NOALLOC_BLOCKS = 1024;
constexpr int NOALLOC_STACK_LDS = 32;
constexpr size_t NOALLOC_SCRATCH_INTS =
    static_cast<size_t>(NOALLOC_BLOCKS) * TRACE_BLOCK_THREADS * (STACK_TOTAL - NOALLOC_STACK_LDS);
static __device__ __align__(256) int scratch[NOALLOC_SCRATCH_INTS];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
int* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch`, looked up once per device.
int noalloc_scratch(int** out) {
  int dev = -1;
  PSDR_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error(PSDR_ERR_ARG, "scene_trace_noalloc: device ordinal beyond the library's per-device table");
    return PSDR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      set_error(PSDR_ERR_HIP, std::string("scene_trace_noalloc: hipGetSymbolAddress of the scratch array failed: ") +
                                  hipGetErrorString(e));
      return PSDR_ERR_HIP;
    }
    g_noalloc_ptr[dev] = static_cast<int*>(p);
  }
  *out = g_noalloc_ptr[dev];
  return PSDR_OK;
}
}  // namespace

extern "C" {

const char* psdr_last_error(void) { return g_last_error.c_str(); }
int psdr_abi_version(void) { return PSDR_ABI_VERSION; }

int psdr_scene_create(const psdr_scene_desc* desc, psdr_scene** out, void* stream) {
  if (!out) {
    set_error(PSDR_ERR_ARG, "out is NULL");
    return PSDR_ERR_ARG;
  }
  *out = nullptr;
  PSDR_TRY(check_desc(desc));  // before anything is allocated or the device is touched
  psdr_scene* s = new (std::nothrow) psdr_scene();
  if (s) {
    s->ws = new (std::nothrow) Workspace();
    if (!s->ws) {
      delete s;
      s = nullptr;
    }
  }
  if (!s) {
    set_error(PSDR_ERR_NOMEM, "host allocation failed");
    return PSDR_ERR_NOMEM;
  }
  const int rc = scene_create_impl(desc, s, as_stream(stream));
  if (rc != PSDR_OK) {
    psdr_scene_destroy(s);
    return rc;
  }
  *out = s;
  return PSDR_OK;
}

int psdr_scene_update(psdr_scene* s, const psdr_scene_desc* desc, void* stream) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "scene is NULL");
    return PSDR_ERR_ARG;
  }
  PSDR_TRY(check_desc(desc));
  CaptureScope cap(as_stream(stream));
  // a boundary prepass still running on the auxiliary streams reads the scene
  PSDR_TRY(aux_join(s, as_stream(stream)));
  PSDR_TRY(scene_set_buffers(s, desc, as_stream(stream)));
  s->ws->invalidate_slots(s);
  return scene_derive(s, as_stream(stream));
}

int psdr_scene_destroy(psdr_scene* s) {
  if (!s) return PSDR_OK;
  (void)hipDeviceSynchronize();
  delete s;  // buffers, streams and events are released by their owners
  return PSDR_OK;
}

int psdr_scene_join(psdr_scene* s, void* stream) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "scene is NULL");
    return PSDR_ERR_ARG;
  }
  CaptureScope cap(as_stream(stream));
  return aux_join(s, as_stream(stream));
}

int psdr_scene_share_workspace(psdr_scene* s, psdr_scene* with, void* stream) {
  if (!s || !with) {
    set_error(PSDR_ERR_ARG, "scene is NULL");
    return PSDR_ERR_ARG;
  }
  if (s->ws == with->ws) return PSDR_OK;
  const hipStream_t st = as_stream(stream);
  CaptureScope cap(st);
  if (capture_active()) {
    set_error(PSDR_ERR_CAPTURE, "share_workspace: buffers cannot change hands under graph capture");
    return PSDR_ERR_CAPTURE;
  }
  PSDR_TRY(aux_join(s, st));
  PSDR_TRY(aux_join(with, st));
  Workspace* old = s->ws;
  if (old->refs == 1) {
    // work queued on any stream may still use its buffers
    PSDR_HIP_CHECK(hipDeviceSynchronize());
    delete old;
  } else {
    old->invalidate_slots(s);
    for (RenderSlot& r : old->slots)
      if (r.owner == s) r.owner = nullptr;
    if (old->last_owner == s) old->last_owner = nullptr;
    --old->refs;
  }
  s->ws = with->ws;
  ++s->ws->refs;
  return PSDR_OK;
}

int psdr_scene_info(const psdr_scene* s, int32_t* n_tris, int32_t* n_edges, int32_t* bvh_depth) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "scene is NULL");
    return PSDR_ERR_ARG;
  }
  if (n_tris) *n_tris = s->n_tris;
  if (n_edges) *n_edges = s->n_edges;
  if (bvh_depth) {
    // depth = longest parent chain (host walk; diagnostic only)
    const int n = s->n_bvh;
    int depth = 0;
    if (n >= 2) {
      std::vector<int> parent(2 * n - 1);
      if (hipMemcpy(parent.data(), s->parent.p, 4ull * (2 * n - 1), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(PSDR_ERR_HIP, "psdr_scene_info: copy failed");
        return PSDR_ERR_HIP;
      }
      std::vector<int> dep(2 * n - 1, -1);
      for (int k = 0; k < 2 * n - 1; ++k) {
        int d = 0, c = k;
        while (c >= 0 && parent[c] >= 0 && d <= 2 * n) {
          c = parent[c];
          ++d;
        }
        depth = std::max(depth, d);
      }
    }
    *bvh_depth = depth;
  }
  return PSDR_OK;
}

int psdr_render_forward(psdr_scene* s, const psdr_config* cfg, const psdr_camera* cam, float* out_rgb, int32_t keep,
                        void* stream) {
  if (!cfg_ok(cfg) || !cam_ok(cam)) {
    set_error(PSDR_ERR_ARG, "render_forward: bad argument");
    return PSDR_ERR_ARG;
  }
  PSDR_TRY(check_render_sizes(*cfg, *cam));
  if (!s || !out_rgb) {
    set_error(PSDR_ERR_ARG, "render_forward: bad argument");
    return PSDR_ERR_ARG;
  }
  CaptureScope cap(as_stream(stream));
  return render_forward(s, *cfg, cam, 1, &out_rgb, keep, as_stream(stream));
}

int psdr_render_forward_multi(psdr_scene* s, const psdr_config* cfg, const psdr_camera* cams, int32_t n_cams,
                              float* const* out_rgb, int32_t keep, void* stream) {
  if (!cfg_ok(cfg) || !cams || n_cams < 1 || n_cams > PSDR_MAX_CAMERAS) {
    set_error(PSDR_ERR_ARG, "render_forward_multi: bad argument (1 to PSDR_MAX_CAMERAS cameras)");
    return PSDR_ERR_ARG;
  }
  for (int c = 0; c < n_cams; ++c) {
    if (!cam_ok(&cams[c])) {
      set_error(PSDR_ERR_ARG, "render_forward_multi: bad camera " + std::to_string(c));
      return PSDR_ERR_ARG;
    }
    PSDR_TRY(check_render_sizes(*cfg, cams[c]));
  }
  if (!s || !out_rgb) {
    set_error(PSDR_ERR_ARG, "render_forward_multi: bad argument");
    return PSDR_ERR_ARG;
  }
  for (int c = 0; c < n_cams; ++c)
    if (!out_rgb[c]) {
      set_error(PSDR_ERR_ARG, "render_forward_multi: output " + std::to_string(c) + " is NULL");
      return PSDR_ERR_ARG;
    }
  CaptureScope cap(as_stream(stream));
  return render_forward(s, *cfg, cams, n_cams, out_rgb, keep, as_stream(stream));
}

int psdr_render_backward(psdr_scene* s, const psdr_config* cfg, const psdr_camera* cam, const float* dL_dI,
                         float* grad_vertices, float* grad_texels, float* grad_emit, float* grad_camera,
                         int32_t terms, void* stream) {
  if (!cfg_ok(cfg) || !cam_ok(cam)) {
    set_error(PSDR_ERR_ARG, "render_backward: bad argument");
    return PSDR_ERR_ARG;
  }
  PSDR_TRY(check_render_sizes(*cfg, *cam));
  if (!s || !dL_dI) {
    set_error(PSDR_ERR_ARG, "render_backward: bad argument");
    return PSDR_ERR_ARG;
  }
  if (!s->ws->find_slot(s, *cfg, *cam)) {
    set_error(PSDR_ERR_STATE, "render_backward: no matching forward render with keep=1");
    return PSDR_ERR_STATE;
  }
  if (terms & ~PSDR_TERM_ALL) {
    set_error(PSDR_ERR_ARG, "render_backward: bad terms mask");
    return PSDR_ERR_ARG;
  }
  CaptureScope cap(as_stream(stream));
  return render_backward(s, *cfg, *cam, dL_dI, grad_vertices, grad_texels, grad_emit, grad_camera, terms,
                         as_stream(stream));
}

namespace {
struct PlainScope {  // sets psdr_scene::plain for one call
  psdr_scene* s;
  explicit PlainScope(psdr_scene* sc) : s(sc) { s->plain = true; }
  ~PlainScope() { s->plain = false; }
};
// the reference launch shape synchronises the host after every trace
int refuse_capture(void* stream, const char* what) {
  CaptureScope cap(as_stream(stream));
  if (!capture_active()) return PSDR_OK;
  set_error(PSDR_ERR_CAPTURE, std::string(what) + ": the plain launch shape synchronises the host and cannot be "
                                                  "captured into a graph");
  return PSDR_ERR_CAPTURE;
}
}  // namespace

int psdr_render_forward_plain(psdr_scene* s, const psdr_config* cfg, const psdr_camera* cam, float* out_rgb,
                              int32_t keep, void* stream) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "render_forward_plain: bad argument");
    return PSDR_ERR_ARG;
  }
  PSDR_TRY(refuse_capture(stream, "render_forward_plain"));
  PlainScope scope(s);
  return psdr_render_forward(s, cfg, cam, out_rgb, keep, stream);
}

int psdr_render_backward_plain(psdr_scene* s, const psdr_config* cfg, const psdr_camera* cam, const float* dL_dI,
                               float* grad_vertices, float* grad_texels, float* grad_emit, float* grad_camera,
                               int32_t terms, void* stream) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "render_backward_plain: bad argument");
    return PSDR_ERR_ARG;
  }
  PSDR_TRY(refuse_capture(stream, "render_backward_plain"));
  PlainScope scope(s);
  return psdr_render_backward(s, cfg, cam, dL_dI, grad_vertices, grad_texels, grad_emit, grad_camera, terms,
                              stream);
}

int psdr_pcg_lane_draws(uint32_t lane0, int32_t n_lanes, int32_t n_draws, uint32_t* out, void* stream) {
  if (n_lanes < 0 || n_draws < 0 || (!out && n_lanes * n_draws > 0)) {
    set_error(PSDR_ERR_ARG, "pcg_lane_draws: bad argument");
    return PSDR_ERR_ARG;
  }
  return lane_draws(lane0, n_lanes, n_draws, out, as_stream(stream));
}

int psdr_camera_rays(const psdr_camera* cam, uint32_t spp, float* o, float* d, int32_t* pix, void* stream) {
  if (!cam_ok(cam) || spp == 0 || !o || !d || !pix) {
    set_error(PSDR_ERR_ARG, "camera_rays: bad argument");
    return PSDR_ERR_ARG;
  }
  return camera_rays(make_camera(*cam), spp, o, d, pix, as_stream(stream));
}

int psdr_camera_param_vjp(const psdr_camera* cam, const float* matrix_adjoints, float* grad) {
  if (!cam_ok(cam) || !matrix_adjoints || !grad) {
    set_error(PSDR_ERR_ARG, "camera_param_vjp: bad argument");
    return PSDR_ERR_ARG;
  }
  camera_param_vjp(*cam, matrix_adjoints, grad);
  return PSDR_OK;
}

int psdr_camera_matrices(const psdr_camera* cam, float* c2w, float* w2c, float* c2r, float* r2c, float* pos) {
  if (!cam_ok(cam)) {
    set_error(PSDR_ERR_ARG, "camera_matrices: bad camera");
    return PSDR_ERR_ARG;
  }
  const CameraK k = make_camera(*cam);
  if (c2w) std::memcpy(c2w, k.c2w.m, 64);
  if (w2c) std::memcpy(w2c, k.w2c.m, 64);
  if (c2r) std::memcpy(c2r, k.c2r.m, 64);
  if (r2c) std::memcpy(r2c, k.r2c.m, 64);
  if (pos) {
    pos[0] = k.pos.x;
    pos[1] = k.pos.y;
    pos[2] = k.pos.z;
  }
  return PSDR_OK;
}

int psdr_scene_trace(psdr_scene* s, const float* ro, const float* rd, int32_t n, int32_t* tri, float* t, float* u,
                     float* v, void* stream) {
  if (!s || n < 0 || n > PSDR_NOALLOC_MAX_RAYS || (n > 0 && (!ro || !rd || !tri || !u || !v))) {
    set_error(PSDR_ERR_ARG, "scene_trace: bad argument (n must lie in [0, (2^31 - 1) / 3])");
    return PSDR_ERR_ARG;
  }
  if (n == 0) return PSDR_OK;
  const int grid = trace_grid(n);
  PSDR_TRY(s->ws->spill.ensure(std::max(s->ws->spill.bytes, spill_bytes(grid))));
  return trace_rays(s->view(), ro, rd, n, nullptr, 1, 0, nullptr, 0, tri, t, u, v, s->ws->spill.as<int>(), grid,
                    as_stream(stream));
}

int psdr_scene_trace_noalloc(psdr_scene* s, const float* ro, const float* rd, int32_t n, int32_t* tri, float* t,
                             float* u, float* v, void* stream) {
  if (!s || n < 0 || (n > 0 && (!ro || !rd || !tri || !u || !v))) {
    set_error(PSDR_ERR_ARG, "scene_trace_noalloc: bad argument");
    return PSDR_ERR_ARG;
  }
  if (n > PSDR_NOALLOC_MAX_RAYS) {
    set_error(PSDR_ERR_ARG, "scene_trace_noalloc: n exceeds PSDR_NOALLOC_MAX_RAYS (ray buffers are indexed as 3 * i "
                            "in 32-bit arithmetic)");
    return PSDR_ERR_ARG;
  }
  // set-up on the first call on this device (also with n == 0): the lookup may
  // load the library's device code, which places the array
  int* spill = nullptr;
  PSDR_TRY(noalloc_scratch(&spill));
  if (n == 0) return PSDR_OK;
  return trace_rays(s->view(), ro, rd, n, nullptr, 1, 0, nullptr, 0, tri, t, u, v, spill, NOALLOC_BLOCKS,
                    as_stream(stream), NOALLOC_STACK_LDS);
}

int psdr_trace_timing(int32_t enable, double* total_ms, int64_t* launches, int64_t* bytes, double* busy_ms) {
  return trace_timing(enable, total_ms, launches, bytes, busy_ms);
}

int psdr_trace_counters(int32_t enable, int64_t* counts, int32_t n_counts) {
  if (!enable && (!counts || n_counts < 0)) {
    set_error(PSDR_ERR_ARG, "trace_counters: counts is NULL");
    return PSDR_ERR_ARG;
  }
  return trace_counters(enable, counts, n_counts);
}

int psdr_get_tuning(psdr_tuning* out) {
  if (!out) {
    set_error(PSDR_ERR_ARG, "get_tuning: output is NULL");
    return PSDR_ERR_ARG;
  }
  const Tuning& t = tuning();
  out->trace_blocks_per_cu = t.trace_bpc.load();
  out->trace_stack_lds = t.trace_stack.load();
  out->shade_blocks_per_cu = t.shade_bpc.load();
  out->adjoint_blocks_per_cu = t.adjoint_bpc.load();
  out->direct_blocks_per_cu = t.direct_bpc.load();
  out->tile_lanes = t.tile_lanes.load();
  return PSDR_OK;
}

int psdr_set_tuning(const psdr_tuning* in) {
  if (!in) {
    set_error(PSDR_ERR_ARG, "set_tuning: input is NULL");
    return PSDR_ERR_ARG;
  }
  const int bpc[4] = {in->trace_blocks_per_cu, in->shade_blocks_per_cu, in->adjoint_blocks_per_cu,
                      in->direct_blocks_per_cu};
  for (int v : bpc)
    if (v < 1 || v > TUNE_BPC_MAX) {
      set_error(PSDR_ERR_ARG, "set_tuning: blocks per CU must lie in 1..64");
      return PSDR_ERR_ARG;
    }
  if (!trace_stack_ok(in->trace_stack_lds)) {
    set_error(PSDR_ERR_ARG, "set_tuning: trace_stack_lds must be 8, 16 or 32");
    return PSDR_ERR_ARG;
  }
  if (in->tile_lanes != 0 && in->tile_lanes < PSDR_MIN_TILE_LANES) {
    set_error(PSDR_ERR_ARG, "set_tuning: tile_lanes must be 0 (no tiling) or at least PSDR_MIN_TILE_LANES");
    return PSDR_ERR_ARG;
  }
  Tuning& t = tuning();
  t.trace_bpc = in->trace_blocks_per_cu;
  t.trace_stack = in->trace_stack_lds;
  t.shade_bpc = in->shade_blocks_per_cu;
  t.adjoint_bpc = in->adjoint_blocks_per_cu;
  t.direct_bpc = in->direct_blocks_per_cu;
  t.tile_lanes = in->tile_lanes;
  return PSDR_OK;
}

int psdr_scene_derived(const psdr_scene* s, float* tri7, int32_t* edge_topo, float* edge_geom, void* stream) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "scene is NULL");
    return PSDR_ERR_ARG;
  }
  hipStream_t st = as_stream(stream);
  if (tri7 && s->n_tris) {
    // tri_geom float4 records -> normal(3), tangent(3), area
    std::vector<float4> g(5ull * s->n_tris);
    PSDR_HIP_CHECK(hipMemcpyAsync(g.data(), s->tri_geom.p, 80ull * s->n_tris, hipMemcpyDeviceToHost, st));
    PSDR_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<float> h(7ull * s->n_tris);
    for (int f = 0; f < s->n_tris; ++f) {
      const float4 n = g[5 * f + 3], tg = g[5 * f + 4];
      float* o = h.data() + 7 * f;
      o[0] = n.x; o[1] = n.y; o[2] = n.z;
      o[3] = tg.x; o[4] = tg.y; o[5] = tg.z;
      o[6] = g[5 * f].w;
    }
    PSDR_HIP_CHECK(hipMemcpyAsync(tri7, h.data(), 4 * h.size(), hipMemcpyHostToDevice, st));
    PSDR_HIP_CHECK(hipStreamSynchronize(st));
  }
  if (edge_topo && s->n_edges)
    PSDR_HIP_CHECK(hipMemcpyAsync(edge_topo, s->edge_topo.p, 24ull * s->n_edges, hipMemcpyDeviceToDevice, st));
  if (edge_geom && s->n_edges)
    PSDR_HIP_CHECK(hipMemcpyAsync(edge_geom, s->edge_geom.p, 64ull * s->n_edges, hipMemcpyDeviceToDevice, st));
  return PSDR_OK;
}

int psdr_scene_samplers(const psdr_scene* s, float* light_pmf, float* light_cmf, float* edge_pmf, float* edge_cmf,
                        float* light_power, float* light_area, void* stream) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "scene is NULL");
    return PSDR_ERR_ARG;
  }
  hipStream_t st = as_stream(stream);
  const hipMemcpyKind k = hipMemcpyDeviceToDevice;
  if (light_pmf && s->n_objects) PSDR_HIP_CHECK(hipMemcpyAsync(light_pmf, s->obj_pmf.p, 4ull * s->n_objects, k, st));
  if (light_cmf && s->n_objects) PSDR_HIP_CHECK(hipMemcpyAsync(light_cmf, s->obj_cmf.p, 4ull * s->n_objects, k, st));
  if (edge_pmf && s->n_edges) PSDR_HIP_CHECK(hipMemcpyAsync(edge_pmf, s->edge_pmf.p, 4ull * s->n_edges, k, st));
  if (edge_cmf && s->n_edges) PSDR_HIP_CHECK(hipMemcpyAsync(edge_cmf, s->edge_cmf.p, 4ull * s->n_edges, k, st));
  if (light_power && s->n_lights)
    PSDR_HIP_CHECK(hipMemcpyAsync(light_power, s->light_power.p, 4ull * s->n_lights, k, st));
  if (light_area && s->n_lights)
    PSDR_HIP_CHECK(hipMemcpyAsync(light_area, s->light_area.p, 4ull * s->n_lights, k, st));
  return PSDR_OK;
}

int psdr_scene_bvh(const psdr_scene* s, uint32_t* morton, int32_t* order, int32_t* child, int32_t* parent, float* box,
                   void* stream) {
  if (!s) {
    set_error(PSDR_ERR_ARG, "scene is NULL");
    return PSDR_ERR_ARG;
  }
  const int n = s->n_bvh;
  if (n < 2) return PSDR_OK;
  hipStream_t st = as_stream(stream);
  const hipMemcpyKind k = hipMemcpyDeviceToDevice;
  if (morton) PSDR_HIP_CHECK(hipMemcpyAsync(morton, s->morton.p, 4ull * n, k, st));
  if (order) PSDR_HIP_CHECK(hipMemcpyAsync(order, s->order.p, 4ull * n, k, st));
  if (child) PSDR_HIP_CHECK(hipMemcpyAsync(child, s->child.p, 8ull * (n - 1), k, st));
  if (parent) PSDR_HIP_CHECK(hipMemcpyAsync(parent, s->parent.p, 4ull * (2 * n - 1), k, st));
  if (box) PSDR_HIP_CHECK(hipMemcpyAsync(box, s->box.p, 48ull * (n - 1), k, st));
  return PSDR_OK;
}

int psdr_last_lane_radiance(const psdr_scene* s, float* out, void* stream) {
  if (!s || !out || s->ws->last_N == 0 || s->ws->last_owner != s) {
    set_error(PSDR_ERR_STATE, "last_lane_radiance: no forward render");
    return PSDR_ERR_STATE;
  }
  PSDR_HIP_CHECK(hipMemcpyAsync(out, s->ws->lane_L.p, 12ull * s->ws->last_N, hipMemcpyDeviceToDevice,
                                as_stream(stream)));
  return PSDR_OK;
}

int psdr_last_render_stats(const psdr_scene* s, int64_t* counts, int32_t n_counts, int64_t* rays_traced) {
  if (!s || s->ws->last_N == 0 || s->ws->last_owner != s) {
    set_error(PSDR_ERR_STATE, "last_render_stats: no forward render");
    return PSDR_ERR_STATE;
  }
  // one block of 128 counters per lane tile: sum them
  const int tiles = std::max(s->ws->last_tiles, 1);
  std::vector<int> raw(128ull * tiles, 0);
  PSDR_HIP_CHECK(hipMemcpy(raw.data(), s->ws->counters_r.p, 4ull * 128 * tiles, hipMemcpyDeviceToHost));
  std::vector<int64_t> c(128, 0);
  for (int t = 0; t < tiles; ++t)
    for (int i = 0; i < 128; ++i) c[i] += raw[128ull * t + i];
  for (int i = 0; i < n_counts && i < 64; ++i) counts[i] = c[i];
  if (rays_traced) {
    // camera rays + per depth d < D: alive paths x (light + BSDF samples, + 1
    // extension ray except at the last depth)
    const Workspace& w = *s->ws;
    const int64_t nLB = static_cast<int64_t>(w.last_cfg.mis_light_samples) + w.last_cfg.mis_bsdf_samples;
    int64_t r = c[0];
    for (int d = 0; d < w.last_D; ++d) r += c[d + 1] * (nLB + (d < w.last_D - 1 ? 1 : 0));
    *rays_traced = r;
  }
  return PSDR_OK;
}

}  // extern "C"
