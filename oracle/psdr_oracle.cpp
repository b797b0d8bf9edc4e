// CPU oracle for the torch_psdr path tracer — TEST INFRASTRUCTURE ONLY (see psdr_oracle.hpp).
// Restates src/scene.cc, src/emitter.cc (AreaLight), src/diffuse_material.cc,
// src/diffuse_bsdf.cc, src/texture.h, src/path_integrator.cc and src/camera.h's Film.
#include "psdr_oracle.hpp"

#include <atomic>
#include <cstdio>
#include <functional>
#include <memory>
#include <numeric>
#include <thread>

using namespace orc;

extern "C" {
// Flat scene description (host arrays).  Objects are meshes first, then light
// meshes, in the order the reference concatenates them (src/scene.cc:8-20).
struct oracle_scene {
  int n_vertices;
  const float* vertices;    // [n_vertices*3]
  const float* d_vertices;  // tangent or NULL
  int n_tris;
  const int* faces;     // [n_tris*3] global vertex indices
  const float* tri_uv;  // [n_tris*6] corner uvs (defaults (0,0),(0,1),(1,0) filled by the caller)
  const int* tri_obj;   // [n_tris]
  int n_objects;
  const int* obj_mat;          // MatType
  const int* obj_albedo;       // [n_objects*3]: float offset, h, w (h=w=0: constant)
  const int* obj_rough;        // [n_objects*3]: 1-channel roughness texture (DiffuseBsdf)
  const int* obj_light;        // light index or -1
  int n_texels;                // floats
  const float* texels;
  const float* d_texels;  // tangent or NULL
  int n_lights;
  const float* light_emit;    // [n_lights*3]
  const float* d_light_emit;  // tangent or NULL
  const int* light_obj;       // [n_lights]
  const int* obj_normal;      // [n_objects*3]: normal map offset (-1: none), h, w; or NULL
};

struct oracle_camera {
  float look_from[3], look_at[3], up[3];
  float vfov;
  int height, width;
  int kind;          // 0 perspective, 1 orthographic
  float view_width;  // orthographic
  float d_params[10];  // tangent direction of look_from, look_at, up, vfov | view_width
};

struct oracle_config {
  uint32_t n_pass, spp_interior, enable_light_visable, spp_primary_edge, spp_secondary_edge, max_bounce,
      mis_light_samples, mis_bsdf_samples, primary_edge_preprocess_rounds, secondary_edge_preprocess_rounds,
      primary_edge_hypercube_resolution, secondary_edge_hypercube_resolution[3];
};
}

namespace {

int n_threads() {
  unsigned h = std::thread::hardware_concurrency();
  return h == 0 ? 1 : static_cast<int>(h);
}
int g_threads = 0;

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
  const int nt = g_threads > 0 ? g_threads : n_threads();
  if (nt <= 1 || n < 1024) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  std::atomic<int64_t> next{0};
  const int64_t block = 256;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&]() {
      for (;;) {
        int64_t b = next.fetch_add(block);
        if (b >= n) break;
        fn(b, std::min(n, b + block));
      }
    });
  for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// Independent CPU BVH (median split); conservative boxes; closest hit with the
// (t, triangle index) order.  Used only to make the oracle fast.
// ---------------------------------------------------------------------------
struct CpuBvh {
  struct Node {
    float lo[3], hi[3];
    int left = -1, right = -1, first = 0, count = 0;
  };
  std::vector<Node> nodes;
  std::vector<int> idx;
  std::vector<V3<float>> p0, e1, e2;
  int n = 0;

  void build(const std::vector<V3<float>>& P0, const std::vector<V3<float>>& P1, const std::vector<V3<float>>& P2) {
    n = static_cast<int>(P0.size());
    p0 = P0;
    e1.resize(n);
    e2.resize(n);
    for (int i = 0; i < n; ++i) {
      e1[i] = P1[i] - P0[i];
      e2[i] = P2[i] - P0[i];
    }
    idx.resize(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<float> blo(3 * n), bhi(3 * n), cen(3 * n);
    // Boxes are padded beyond the triangle by more than the rounding of the
    // exact test, which can accept a ray passing outside the triangle by a few
    // ulp of the coordinates involved (rays through vertices): relative to the
    // box's own coordinates and to the scene's extent.
    float scale = 0.f;
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a)
        scale = std::max(scale, std::max(std::fabs(P0[i][a]), std::max(std::fabs(P1[i][a]), std::fabs(P2[i][a]))));
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        float x0 = P0[i][a], x1 = P1[i][a], x2 = P2[i][a];
        float lo = std::min(x0, std::min(x1, x2)), hi = std::max(x0, std::max(x1, x2));
        float pad = (std::fabs(lo) + std::fabs(hi)) * 1e-5f + scale * 1e-6f + 1e-30f;
        blo[3 * i + a] = lo - pad;
        bhi[3 * i + a] = hi + pad;
        cen[3 * i + a] = 0.5f * (lo + hi);
      }
    nodes.clear();
    nodes.reserve(2 * std::max(n, 1));
    std::function<int(int, int)> rec = [&](int b, int e) -> int {
      Node nd;
      for (int a = 0; a < 3; ++a) {
        nd.lo[a] = std::numeric_limits<float>::infinity();
        nd.hi[a] = -std::numeric_limits<float>::infinity();
      }
      for (int k = b; k < e; ++k)
        for (int a = 0; a < 3; ++a) {
          nd.lo[a] = std::min(nd.lo[a], blo[3 * idx[k] + a]);
          nd.hi[a] = std::max(nd.hi[a], bhi[3 * idx[k] + a]);
        }
      const int me = static_cast<int>(nodes.size());
      nodes.push_back(nd);
      if (e - b <= 4) {
        nodes[me].first = b;
        nodes[me].count = e - b;
        return me;
      }
      int ax = 0;
      float ext = -1.f;
      for (int a = 0; a < 3; ++a)
        if (nd.hi[a] - nd.lo[a] > ext) {
          ext = nd.hi[a] - nd.lo[a];
          ax = a;
        }
      const int m = (b + e) / 2;
      std::nth_element(idx.begin() + b, idx.begin() + m, idx.begin() + e,
                       [&](int x, int y) { return cen[3 * x + ax] < cen[3 * y + ax]; });
      const int l = rec(b, m);
      const int r = rec(m, e);
      nodes[me].left = l;
      nodes[me].right = r;
      return me;
    };
    if (n > 0) rec(0, n);
  }

  static bool box_hit(const Node& nd, const V3<float>& o, const V3<float>& inv, float tbest) {
    float t0 = 0.f, t1 = std::numeric_limits<float>::infinity();
    for (int a = 0; a < 3; ++a) {
      float ta = (nd.lo[a] - o[a]) * inv[a];
      float tb = (nd.hi[a] - o[a]) * inv[a];
      float tn = std::fmin(ta, tb), tf = std::fmax(ta, tb);
      if (std::isnan(tn)) tn = -std::numeric_limits<float>::infinity();
      if (std::isnan(tf)) tf = std::numeric_limits<float>::infinity();
      t0 = std::max(t0, tn);
      t1 = std::min(t1, tf);
    }
    t1 = t1 * 1.0000004f;
    return t0 <= t1 && t0 <= tbest * 1.0000004f;
  }

  Hit closest(const V3<float>& o, const V3<float>& d) const {
    Hit h;
    if (n == 0) return h;
    const V3<float> inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node& nd = nodes[stack[--sp]];
      if (!box_hit(nd, o, inv, h.t)) continue;
      if (nd.left < 0) {
        for (int k = nd.first; k < nd.first + nd.count; ++k) {
          const int tri = idx[k];
          float t, u, v;
          if (ray_tri(o, d, p0[tri], e1[tri], e2[tri], &t, &u, &v)) {
            if (t < h.t || (t == h.t && tri < h.tri)) {
              h.t = t;
              h.tri = tri;
              h.u = u;
              h.v = v;
            }
          }
        }
      } else {
        stack[sp++] = nd.left;
        stack[sp++] = nd.right;
      }
    }
    return h;
  }

  // Closest hit plus the second-smallest t among the other triangles the ray
  // hits (boxes pruned against that second distance).
  Hit closest2(const V3<float>& o, const V3<float>& d, float* t2) const {
    Hit h;
    *t2 = std::numeric_limits<float>::infinity();
    if (n == 0) return h;
    const V3<float> inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node& nd = nodes[stack[--sp]];
      if (!box_hit(nd, o, inv, *t2)) continue;
      if (nd.left < 0) {
        for (int k = nd.first; k < nd.first + nd.count; ++k) {
          const int tri = idx[k];
          float t, u, v;
          if (!ray_tri(o, d, p0[tri], e1[tri], e2[tri], &t, &u, &v)) continue;
          if (t < h.t || (t == h.t && tri < h.tri)) {
            *t2 = std::min(*t2, h.t);
            h.t = t;
            h.tri = tri;
            h.u = u;
            h.v = v;
          } else {
            *t2 = std::min(*t2, t);
          }
        }
      } else {
        stack[sp++] = nd.left;
        stack[sp++] = nd.right;
      }
    }
    return h;
  }

  Hit brute(const V3<float>& o, const V3<float>& d) const {
    Hit h;
    for (int tri = 0; tri < n; ++tri) {
      float t, u, v;
      if (ray_tri(o, d, p0[tri], e1[tri], e2[tri], &t, &u, &v)) {
        if (t < h.t || (t == h.t && tri < h.tri)) {
          h.t = t;
          h.tri = tri;
          h.u = u;
          h.v = v;
        }
      }
    }
    return h;
  }
};

// ---------------------------------------------------------------------------
// Intersection record (src/intersection.h, src/scene.cc:106-161)
// ---------------------------------------------------------------------------
// SH9Light::eval_impl_ (src/emitter.cc:251-279): sum over k of C_k Y_k(w), the
// real spherical harmonics up to degree 2 with fp32 constants sqrt(1/(4 pi)),
// sqrt(3/(4 pi)), sqrt(15/(4 pi)), sqrt(5/(16 pi)), sqrt(15/(16 pi)).
template <class F> V3<F> sh9_radiance(const F* C, const V3<F>& w) {
  const float c0 = std::sqrt(REAL_INV_4PI), c1 = std::sqrt(0.75f * REAL_INV_PI);
  const float c2 = std::sqrt(3.75f * REAL_INV_PI), c3 = std::sqrt(0.3125f * REAL_INV_PI);
  const float c4 = std::sqrt(0.9375f * REAL_INV_PI);
  const F xx = w.x * w.x, yy = w.y * w.y, zz = w.z * w.z;
  const F xy = w.x * w.y, yz = w.y * w.z, zx = w.z * w.x;
  const F Y[9] = {F(c0),      F(c1) * w.y, F(c1) * w.z, F(c1) * w.x, F(c2) * xy, F(c2) * yz,
                  F(c3) * ((F(2.0f) * zz - xx) - yy), F(c2) * zx, F(c4) * (xx - yy)};
  V3<F> r = V3<F>{C[0], C[1], C[2]} * Y[0];
  for (int k = 1; k < 9; ++k) r = r + V3<F>{C[3 * k], C[3 * k + 1], C[3 * k + 2]} * Y[k];
  return r;
}

template <class F> struct Its {
  bool valid = false, is_light = false;
  V3<F> p, n, t, wi;
  F d2 = F(0.f);
  V2 uv;
  F jac = F(1.f);
  int obj = -1, tri = -1, light = -1;
};

// DiffuseBsdfMaterial::eval_ (src/diffuse_bsdf.cc:57-89) for shading normal N,
// direction to the viewer V and to the light L: color (a + b t) with
// a = 1 / (pi + (3 pi - 4) / 6 sigma), b = sigma a, t = V.L - max(N.L, 0)
// max(N.V, 0), and t replaced by max(N.L, N.V) + REAL_EPS where it is positive.
template <class F>
V3<F> diffuse_bsdf_value(const V3<F>& color, const F& sigma, const V3<F>& N, const V3<F>& Vw, const V3<F>& L) {
  const F div = F(1.0f) / (F(REAL_PI) + F((3.0f * REAL_PI - 4.0f) / 6.0f) * sigma);
  const F a = div;
  const F b = sigma * div;
  F NdotL = dot(N, L);
  if (val(NdotL) < 0.f) NdotL = F(0.f);
  F NdotV = dot(N, Vw);
  if (val(NdotV) < 0.f) NdotV = F(0.f);
  const F VdotL = dot(Vw, L);
  F t = VdotL - NdotL * NdotV;
  if (val(t) > 0.f) t = (val(NdotL) >= val(NdotV) ? NdotL : NdotV) + F(REAL_EPS);
  return color * (a + b * t);
}

template <class F> struct Scene {
  std::vector<V3<F>> V;
  std::vector<Tri<F>> tris;
  std::vector<Edge<F>> edges;
  std::vector<Object> objs;
  std::vector<Light<F>> lights;
  std::vector<F> texels;
  Discrete light_sampler;  // over objects, weight = light power (0 for non-lights)
  Discrete edge_sampler;   // over all edges, weight = length
  float lights_total = 0.f;  // sum of light powers (src/scene.cc:59-60)
  int n_edge_tris = 0;       // triangles whose edges are sampled (all but an environment cube)
  CpuBvh bvh;

  void build(const oracle_scene& s, bool use_tangent) {
    V.resize(s.n_vertices);
    for (int i = 0; i < s.n_vertices; ++i)
      for (int a = 0; a < 3; ++a) {
        float dv = (use_tangent && s.d_vertices) ? s.d_vertices[3 * i + a] : 0.f;
        set(V[i][a], s.vertices[3 * i + a], dv);
      }
    texels.resize(s.n_texels);
    for (int i = 0; i < s.n_texels; ++i)
      set(texels[i], s.texels[i], (use_tangent && s.d_texels) ? s.d_texels[i] : 0.f);
    objs.resize(s.n_objects);
    for (int o = 0; o < s.n_objects; ++o) {
      objs[o].mat = s.obj_mat[o];
      objs[o].albedo = {s.obj_albedo[3 * o], s.obj_albedo[3 * o + 1], s.obj_albedo[3 * o + 2], 3};
      objs[o].rough = {s.obj_rough[3 * o], s.obj_rough[3 * o + 1], s.obj_rough[3 * o + 2], 1};
      if (s.obj_normal && s.obj_normal[3 * o] >= 0 && objs[o].mat == MAT_DIFFUSE_BSDF) {
        objs[o].normal = {s.obj_normal[3 * o], s.obj_normal[3 * o + 1], s.obj_normal[3 * o + 2], 3};
        objs[o].has_normal = true;
      }
      objs[o].light = s.obj_light[o];
      objs[o].tri_begin = s.n_tris;
      objs[o].tri_count = 0;
    }
    tris.resize(s.n_tris);
    std::vector<V3<float>> P0(s.n_tris), P1(s.n_tris), P2(s.n_tris);
    for (int f = 0; f < s.n_tris; ++f) {
      const int* fi = s.faces + 3 * f;
      const float* uv = s.tri_uv + 6 * f;
      tris[f] = make_tri<F>(V[fi[0]], V[fi[1]], V[fi[2]], V2{uv[0], uv[1]}, V2{uv[2], uv[3]}, V2{uv[4], uv[5]});
      tris[f].obj = s.tri_obj[f];
      for (int k = 0; k < 3; ++k) tris[f].vi[k] = fi[k];
      Object& ob = objs[s.tri_obj[f]];
      ob.tri_begin = std::min(ob.tri_begin, f);
      ob.tri_count++;
      P0[f] = detach(tris[f].p0);
      P1[f] = detach(tris[f].p1);
      P2[f] = detach(tris[f].p2);
    }
    bvh.build(P0, P1, P2);
    // lights (src/emitter.cc:4-14): power = 2pi * luminance(emit) * area
    lights.resize(s.n_lights);
    std::vector<float> obj_power(s.n_objects, 0.f);
    for (int l = 0; l < s.n_lights; ++l) {
      Light<F>& L = lights[l];
      for (int a = 0; a < 3; ++a)
        set(L.emit[a], s.light_emit[3 * l + a], (use_tangent && s.d_light_emit) ? s.d_light_emit[3 * l + a] : 0.f);
      L.obj = s.light_obj[l];
      const Object& ob = objs[L.obj];
      std::vector<float> areas(ob.tri_count);
      for (int k = 0; k < ob.tri_count; ++k) areas[k] = detach(tris[ob.tri_begin + k].area);
      L.total_area = chunked_sum(areas.data(), ob.tri_count);
      L.tri_sampler.build(areas.data(), ob.tri_count);
      const V3<float> e = detach(L.emit);
      const float lum = (e.x * 0.2126f + e.y * 0.7152f) + e.z * 0.0722f;
      L.power = (REAL_2PI * lum) * L.total_area;
      if (ob.mat == MAT_ENV_MAP || ob.mat == MAT_ENV_SH9) L.power = REAL_4PI;  // src/emitter.cc:129, 247
      obj_power[L.obj] = L.power;
    }
    light_sampler.build(obj_power.data(), s.n_objects);
    lights_total = chunked_sum(obj_power.data(), s.n_objects);
    // edges of every mesh, light meshes included, not the environment cube
    // (src/scene.cc:30, 80-101)
    n_edge_tris = s.n_tris;
    if (s.n_objects > 0 && (objs.back().mat == MAT_ENV_MAP || objs.back().mat == MAT_ENV_SH9))
      n_edge_tris = objs.back().tri_begin;
    std::vector<EdgeTopo> topo = edge_topology(s.faces, n_edge_tris);
    edges.resize(topo.size());
    std::vector<float> lens(topo.size());
    for (size_t e = 0; e < topo.size(); ++e) {
      edges[e] = make_edge<F>(topo[e], V.data());
      lens[e] = detach(edges[e].len);
    }
    edge_sampler.build(lens.data(), static_cast<int>(lens.size()));
  }
  static void set(float& x, float v, float) { x = v; }
  static void set(Dual& x, float v, float d) { x = Dual(v, d); }

  // Light::eval of light l at its hit: AreaLight's two-sided constant
  // (src/emitter.cc:16-23); EnvLight's cube map at the uv (src/emitter.cc:133-140
  // through DiffuseLightMaterial::emit, src/diffuse_material.h:53-58);
  // SH9Light's sum over k of C_k Y_k(wi) (src/emitter.cc:251-279).
  V3<F> radiance(int l, const Its<F>& it) const {
    const Light<F>& L = lights[l];
    const Object& ob = objs[L.obj];
    if (ob.mat == MAT_ENV_MAP) return tex3(ob.albedo, it.uv, texels);
    if (ob.mat == MAT_ENV_SH9) return sh9_radiance<F>(&texels[ob.albedo.offset], it.wi);
    return L.emit;
  }

  // Texture lookup (src/texture.h:24-74): nearest texel, rows flipped by the
  // Python wrapper, so image row = H-1-v0.
  template <class G> V3<G> tex3(const Texture& tx, V2 uv, const std::vector<G>& tex) const {
    if (tx.h == 0) return {tex[tx.offset], tex[tx.offset + 1], tex[tx.offset + 2]};
    const int idx = texel_index(tx, uv);
    return {tex[tx.offset + 3 * idx], tex[tx.offset + 3 * idx + 1], tex[tx.offset + 3 * idx + 2]};
  }
  template <class G> G tex1(const Texture& tx, V2 uv, const std::vector<G>& tex) const {
    if (tx.h == 0) return tex[tx.offset];
    return tex[tx.offset + texel_index(tx, uv)];
  }
  static int texel_index(const Texture& tx, V2 uv) {
    const float u = uv.x * static_cast<float>(tx.w) - 0.5f;
    const float v = uv.y * static_cast<float>(tx.h) - 0.5f;
    const float u0 = std::fmin(std::fmax(std::rint(u), 0.f), static_cast<float>(tx.w - 1));
    const float v0 = std::fmin(std::fmax(std::rint(v), 0.f), static_cast<float>(tx.h - 1));
    const int iu = static_cast<int>(u0), iv = static_cast<int>(v0);
    return (tx.h - 1 - iv) * tx.w + iu;
  }

  // Scene::hit reconstruction; `O` is the (differentiable) ray origin.
  Its<F> reconstruct(const Hit& h, const V3<F>& O, bool valid_in) const {
    Its<F> it;
    it.valid = valid_in && h.tri >= 0;
    if (!it.valid) return it;
    const Tri<F>& T = tris[h.tri];
    const float u = h.u, v = h.v;
    it.tri = h.tri;
    it.obj = T.obj;
    it.light = objs[T.obj].light;
    it.is_light = it.light >= 0;
    it.p = (T.p0 + F(u) * (T.p1 - T.p0)) + F(v) * (T.p2 - T.p0);
    it.uv.x = (T.uv0.x + u * (T.uv1.x - T.uv0.x)) + v * (T.uv2.x - T.uv0.x);
    it.uv.y = (T.uv0.y + u * (T.uv1.y - T.uv0.y)) + v * (T.uv2.y - T.uv0.y);
    const V3<F> wu = it.p - O;
    it.wi = -normalize(wu);
    it.d2 = squared_norm(wu);
    const bool front = val(dot(it.wi, T.n)) > 0.f;
    it.n = front ? T.n : -T.n;
    it.t = front ? T.t : -T.t;
    it.jac = T.area / F(detach(T.area));
    return it;
  }
  Its<F> hit(const V3<F>& O, const V3<F>& D, bool valid) const {
    if (!valid) return Its<F>{};
    const Hit h = bvh.closest(detach(O), detach(D));
    return reconstruct(h, O, valid);
  }

  // Materials --------------------------------------------------------------
  struct Sampled {
    V3<F> wo;
    float pdf = 0.f;
    V3<F> bsdf;
    bool valid = false;
  };
  // DiffuseBsdf shading normal (src/diffuse_bsdf.cc:13-19, 61-68): the
  // geometric normal, or with a normal map normalize(TBN (2 n - 1)) where TBN
  // has the columns T, B = Ng x T, Ng (the matrix-vector product summed in that
  // column order).
  V3<F> shading_normal(const Object& ob, const Its<F>& it) const {
    if (!ob.has_normal) return it.n;
    const V3<F> tn = tex3(ob.normal, it.uv, texels);
    const V3<F> v = tn * F(2.0f) - V3<F>(F(1.0f));
    const V3<F> B = cross(it.n, it.t);
    return normalize((it.t * v.x + B * v.y) + it.n * v.z);
  }
  // DiffuseBsdf shading value (src/diffuse_bsdf.cc:42-75)
  V3<F> bsdf_diffuse_bsdf(const Object& ob, const Its<F>& it, const V3<F>& L) const {
    return diffuse_bsdf_value(tex3(ob.albedo, it.uv, texels), tex1(ob.rough, it.uv, texels), shading_normal(ob, it),
                              it.wi, L);
  }
  // Material::eval (diffuse_material.cc:29-34 / diffuse_bsdf.cc:42-75)
  V3<F> eval(const Its<F>& it, const V3<F>& wo) const {
    const Object& ob = objs[it.obj];
    if (ob.mat == MAT_DIFFUSE_BSDF) return bsdf_diffuse_bsdf(ob, it, wo);
    return tex3(ob.albedo, it.uv, texels) * F(REAL_INV_PI);
  }
  // Material::pdf
  F pdf(const Its<F>& it, const V3<F>& wo) const { return dot(it.n, wo) * F(REAL_INV_PI); }
  // Material::sample (diffuse_material.cc:3-27, diffuse_bsdf.cc:3-40)
  Sampled sample(const Its<F>& it, float ru, float rv, bool valid) const {
    Sampled s;
    const Object& ob = objs[it.obj];
    const V3<F>& T = it.t;
    const V3<F> B = cross(it.n, T);
    const V3<F> N = ob.mat == MAT_DIFFUSE_BSDF ? shading_normal(ob, it) : it.n;
    const float cos_t = safe_sqrt(ru);
    const float sin_t = safe_sqrt(1.0f - ru);
    const float phi = rv * REAL_2PI;
    float sp, cp;
    psdr_sincos(phi, &sp, &cp);
    const float u = cp * sin_t;
    const float v = sp * sin_t;
    s.wo = normalize((F(u) * T + F(v) * B) + F(cos_t) * N);
    s.pdf = cos_t * REAL_INV_PI;
    s.valid = valid && it.valid;
    if (ob.mat == MAT_DIFFUSE_BSDF) {
      s.valid = s.valid && val(dot(it.n, it.wi)) > 0.f;
      s.bsdf = bsdf_diffuse_bsdf(ob, it, s.wo);
    } else {
      s.bsdf = tex3(ob.albedo, it.uv, texels) * F(REAL_INV_PI);
    }
    return s;
  }

  // AreaLight::sample_light_point_ (src/emitter.cc:30-59)
  V3<F> sample_light_point(int l, float ru, float rv, float* pdf) const {
    const Light<F>& L = lights[l];
    float pt;
    const int k = L.tri_sampler.sample_reuse(ru, &pt);
    const Tri<F>& T = tris[objs[L.obj].tri_begin + k];
    const float sv = safe_sqrt(rv);
    const float u = ru * sv;
    const float v = 1.0f - sv;
    *pdf = 1.0f / L.total_area;
    return (T.p0 + (T.p1 - T.p0) * F(u)) + (T.p2 - T.p0) * F(v);
  }
};

inline float mis_weight(float w1, float w2, uint32_t n1, uint32_t n2) {
  const float a = static_cast<float>(n1) * w1;
  const float b = static_cast<float>(n2) * w2;
  const float sa = a * a, sb = b * b;
  return sa / (sa + sb);
}

// PathIntegrator::trace2light (src/path_integrator.cc:80-186).
// The last depth still draws the extension sample (the RNG stream advances);
// its ray cannot change L, so it is not traced.
template <class F>
V3<F> trace2light(const Scene<F>& sc, Pcg32& rng, const Its<F>& its0, bool valid_, uint32_t nL, uint32_t nB,
                  uint32_t D, bool elv) {
  V3<F> T(F(1.f)), L(F(0.f));
  Its<F> its = its0;
  bool valid = valid_ && its0.valid;
  for (uint32_t depth = 0; depth < D; ++depth) {
    {
      const bool hit_light = valid && its.is_light;
      if (elv && depth == 0 && hit_light) L = L + T * sc.radiance(its.light, its);
      valid = valid && !hit_light;
    }
    for (uint32_t i = 0; i < nL; ++i) {
      float ru = rng.next_float(valid);
      float rv = rng.next_float(valid);
      if (!valid) continue;
      float pdf_choose;
      const int obj = sc.light_sampler.sample_reuse(ru, &pdf_choose);
      const int l = sc.objs[obj].light;
      if (l < 0) continue;
      float pdf_pt;
      const V3<F> lp = sc.sample_light_point(l, ru, rv, &pdf_pt);
      const float pdf_light = pdf_pt * pdf_choose;
      const V3<F> dir = normalize(lp - its.p);
      const Its<F> itl = sc.hit(its.p, dir, valid);
      const bool is_hit = itl.valid && itl.is_light && squared_norm(detach(itl.p) - detach(lp)) < REAL_EPS;
      if (!is_hit) continue;
      const V3<F> emit = sc.radiance(l, itl);
      const V3<F> bsdf = sc.eval(its, dir);
      const F tmp = dot(itl.n, -dir) / itl.d2;
      const F G = dot(its.n, dir) * tmp;
      const V3<F> value = (((emit * G) * bsdf) * itl.jac) / F(pdf_light);
      float weight = 1.0f / static_cast<float>(nL);
      if (nB > 0) {
        const float pdf_bsdf = detach(sc.pdf(its, dir) * tmp);
        weight = weight * mis_weight(pdf_light, pdf_bsdf, nL, nB);
      }
      L = L + (T * F(weight)) * value;
    }
    for (uint32_t i = 0; i < nB; ++i) {
      const float ru = rng.next_float(valid);
      const float rv = rng.next_float(valid);
      if (!valid) continue;
      const auto smp = sc.sample(its, ru, rv, valid);
      const Its<F> itl = sc.hit(its.p, smp.wo, valid);
      const bool is_hit = valid && itl.valid && itl.is_light;
      if (!is_hit) continue;
      const V3<F> emit = sc.radiance(itl.light, itl);
      const F tmp = dot(itl.n, -smp.wo) / itl.d2;
      const F G = dot(its.n, smp.wo) * tmp;
      const float pdf_bsdf = detach(F(smp.pdf) * tmp);
      const V3<F> value = (((emit * G) * smp.bsdf) * itl.jac) / F(pdf_bsdf);
      float weight = 1.0f / static_cast<float>(nB);
      if (nL > 0) {
        const Light<F>& LL = sc.lights[itl.light];
        const float pdf_light = (LL.power / sc.lights_total) * (1.0f / LL.total_area);
        weight = weight * mis_weight(pdf_bsdf, pdf_light, nB, nL);
      }
      L = L + (T * F(weight)) * value;
    }
    {
      const float ru = rng.next_float(valid);
      const float rv = rng.next_float(valid);
      if (depth + 1 == D) break;
      if (!valid) continue;
      const auto smp = sc.sample(its, ru, rv, valid);
      const Its<F> itn = sc.hit(its.p, smp.wo, valid);
      valid = valid && smp.valid && itn.valid;
      if (valid) {
        const F tmp = dot(itn.n, -smp.wo) / itn.d2;
        const F G = dot(its.n, smp.wo) * tmp;
        const float pdf = detach(F(smp.pdf) * tmp);
        T = T * (((smp.bsdf * itn.jac) * G) / F(pdf));
      }
      its = itn;
    }
  }
  if (!std::isfinite(val(squared_norm(L)))) L = V3<F>(F(0.f));
  return L;
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// Boundary estimators (src/path_integrator.cc:11-78, 188-353)
// ---------------------------------------------------------------------------
template <class F> V3<F> hit_triangle(const V3<float>& o, const V3<F>& d, const V3<float>& p0, const V3<float>& e1,
                                      const V3<float>& e2) {
  const V3<float> T = o - p0;
  const V3<F> P = cross(d, lift<F>(e2));
  const V3<float> Q = cross(T, e1);
  const F inv = F(1.0f) / clamp_(dot(P, lift<F>(e1)), REAL_EPS, 1.0f - REAL_EPS);
  const F u = dot(P, lift<F>(T)) * inv;
  const F v = dot(lift<F>(Q), d) * inv;
  return (lift<F>(p0) + u * lift<F>(e1)) + v * lift<F>(e2);
}

// Test-only switch (see oracle_set_direct_unit_normal): the direct boundary
// term with the shadow line's direction and normal on the receiver normalised
// and the shadow point's motion projected on that normal by a dot product.
bool g_direct_unit = false;

template <class F>
V3<F> boundary_term(const V3<F>& xs, const V3<F>& xd, const V3<F>& xb, const V3<F>& ns, const V3<F>& nd,
                    const Edge<F>& E, bool ad) {
  const V3<float> xs_ = detach(xs), xd_ = detach(xd), xb_ = detach(xb), ns_ = detach(ns), nd_ = detach(nd);
  const V3<F> dir = normalize(xb - lift<F>(xs_));
  const V3<float> dir_ = detach(dir);
  V3<float> ev = normalize(detach(E.p1) - detach(E.p0));
  if (dot(cross(dir_, ev), detach(E.out)) < 0.f) ev = -ev;
  const V3<float> n_plane = cross(dir_, ev);
  V3<float> edge_proj = cross(nd_, n_plane);
  if (g_direct_unit) edge_proj = normalize(edge_proj);
  const float cos_s = std::fabs(dot(ns_, dir_));
  const float cos_b = dot(ev, dir_);
  const float cos_d = dot(edge_proj, dir_);
  const float sin_b = safe_sqrt(1.0f - cos_b * cos_b);
  const float sin_d = safe_sqrt(1.0f - cos_d * cos_d);
  const float len_sb = norm(xb_ - xs_);
  const float len_sd = norm(xd_ - xs_);
  const float Jb = ((len_sd * len_sb) * sin_b) / (sin_d * cos_s);
  const float dG = (cos_s * dot(nd_, -dir_)) / (len_sd * len_sd);
  V3<F> fb;
  if (ad) {
    const V3<float> nep = cross(nd_, edge_proj);
    V3<F> xh = hit_triangle<F>(xs_, dir, xd_, edge_proj, nep);
    xh = xh - lift<F>(detach(xh));
    if (g_direct_unit)
      fb = V3<F>(dot(xh, lift<F>(nep)) * F(dG));
    else
      fb = xh * lift<F>(nep * dG);
  } else {
    fb = V3<F>(F(dG));
  }
  return fb * F(Jb);
}

struct BoundaryOut {
  int px = 0, py = 0;
  bool valid = false;
};

// Test-only switch: rescale each boundary sample from the reference's
// normalisation (world measure, samples summed per pixel and divided by the
// edge spp) to the raster measure of a pixel mean (divided by all spp*H*W
// samples), so the boundary terms can be compared with finite differences.
bool g_fd_measure = false;

// eval_primary_boundary (src/path_integrator.cc:202-251), with the
// camera-origin term (o - detach o) . (n (x) [1, 1, -1]) of :240-244.
template <class F>
V3<F> eval_primary(const Scene<F>& sc, const Scene<float>& scv, const Camera& cam, Pcg32& rng, float s1, uint32_t nL,
                   uint32_t D, bool ad, BoundaryOut* bo) {
  bo->valid = false;
  float u = s1, pdf_e;
  if (sc.edges.empty()) return V3<F>(F(0.f));
  const int e = sc.edge_sampler.sample_reuse(u, &pdf_e);
  const Edge<F>& E = sc.edges[e];
  const V3<F> ep = E.p0 + F(u) * (E.p1 - E.p0);
  const F pdf_ep = F(pdf_e) / E.len;
  const V3<float> epv = detach(ep);
  bool valid = cam.ray_with_point(epv, &bo->px, &bo->py);
  V3<float> o, d;
  V3<Dual> o_ad;
  cam.point_ray(epv, V3<float>{tan_(ep.x), tan_(ep.y), tan_(ep.z)}, &o, &d, &o_ad);
  const V3<float> n = cross(d, normalize(detach(E.p1) - detach(E.p0)));
  const V3<float> da = d - n * REAL_EPS;
  const V3<float> db = d + n * REAL_EPS;
  const Its<float> ia = scv.hit(o, da, valid);
  const Its<float> ib = scv.hit(o, db, valid);
  valid = valid && ia.valid && ib.valid &&
          (norm(ia.p - epv) < 10.0f * REAL_EPS || norm(ib.p - epv) < 10.0f * REAL_EPS);
  const V3<float> La = trace2light<float>(scv, rng, ia, valid, nL, 1, D, false);
  const V3<float> Lb = trace2light<float>(scv, rng, ib, valid, nL, 1, D, false);
  const V3<float> dL = La - Lb;
  V3<F> value;
  if (ad) {
    const V3<float> nm{n.x, n.y, -n.z};
    const F xn = dot(ep - lift<F>(epv), lift<F>(n)) + dot(origin_as<F>(o, o_ad) - lift<F>(o), lift<F>(nm));
    value = (lift<F>(dL) * xn) / F(detach(pdf_ep));
    if (g_fd_measure && valid)
      value = value * F(static_cast<float>(cam.raster_jacobian(epv, d) / (static_cast<double>(cam.H) * cam.W)));
  } else {
    value = lift<F>(dL) / pdf_ep;
  }
  valid = valid && std::isfinite(val(squared_norm(value)));
  bo->valid = valid;
  return value;
}

// eval_direct_boundary (src/path_integrator.cc:287-329)
template <class F>
V3<F> eval_direct(const Scene<F>& sc, const Camera& cam, float sx, float sy, float sz, bool ad, BoundaryOut* bo) {
  bo->valid = false;
  if (sc.edges.empty()) return V3<F>(F(0.f));
  float ue = sx, pdf_e;
  const int e = sc.edge_sampler.sample_reuse(ue, &pdf_e);
  const Edge<F>& E = sc.edges[e];
  const V3<F> ep = E.p0 + F(ue) * (E.p1 - E.p0);
  const F pdf_ep = F(pdf_e) / E.len;
  float y = sy, pdf_choose;
  const int obj = sc.light_sampler.sample_reuse(y, &pdf_choose);
  const int l = sc.objs[obj].light;
  if (l < 0) return V3<F>(F(0.f));
  float pdf_light;
  const V3<F> lp = sc.sample_light_point(l, y, sz, &pdf_light);
  pdf_light = pdf_light * pdf_choose;
  const V3<F> dir = normalize(ep - lp);
  const V3<float> dv = detach(dir);
  bool valid = (dot(detach(E.nl), dv) * dot(detach(E.nr), dv) < 0.f) && norm(detach(E.nl) - detach(E.nr)) > REAL_EPS;
  const Its<F> is0 = sc.hit(ep, -dir, valid);
  valid = valid && is0.valid && is0.is_light && squared_norm(detach(is0.p) - detach(lp)) < REAL_EPS;
  const Its<F> id0 = sc.hit(ep, dir, valid);
  valid = valid && id0.valid && !id0.is_light;
  int px = 0, py = 0;
  const V3<float> xd = detach(id0.p);
  const bool vp = valid ? cam.ray_with_point(xd, &px, &py) : false;
  valid = valid && vp;
  V3<float> co, cd;
  V3<Dual> co_ad;
  cam.point_ray(xd, V3<float>{0.f, 0.f, 0.f}, &co, &cd, &co_ad);
  const Its<F> i2 = sc.hit(lift<F>(co), lift<F>(cd), valid);
  valid = valid && i2.valid && !i2.is_light && squared_norm(detach(i2.p) - xd) < REAL_EPS;
  bo->px = px;
  bo->py = py;
  if (!valid) return V3<F>(F(0.f));
  const F pdf_xw = ((is0.d2 * pdf_ep) * F(pdf_light)) / dot(is0.n, dir);
  const V3<F> Tb = boundary_term<F>(is0.p, id0.p, ep, is0.n, id0.n, E, ad) / pdf_xw;
  const V3<float> Ts = detach(sc.radiance(is0.light, is0));
  Its<F> id0c = id0;
  const V3<float> Td = detach(sc.eval(id0c, lift<F>(-cd)));
  V3<F> value = (Tb * lift<F>(Ts)) * lift<F>(Td);
  if (g_fd_measure && ad) {
    // receiver area -> area perpendicular to the camera ray -> raster
    const double cosr = std::fabs(static_cast<double>(dot(detach(id0.n), cd)));
    value = value * F(static_cast<float>(cam.raster_jacobian(xd, cd) * cosr /
                                         (static_cast<double>(cam.H) * cam.W)));
  }
  valid = valid && std::isfinite(val(squared_norm(value)));
  bo->valid = valid;
  return value;
}

}  // namespace

namespace {

// HyperCubeSampler (src/sampler.h:201-255): row-major cells, dim 0 slowest;
// the last coordinate picks a cell through a DiscreteSampler over its mass.
struct HyperCube {
  int dims = 1;
  int res[3] = {1, 1, 1};
  int cells = 1;
  float unit[3] = {1.f, 1.f, 1.f};
  Discrete pick;
  void init(int d, const uint32_t* r) {
    dims = d;
    cells = 1;
    for (int i = 0; i < d; ++i) {
      res[i] = static_cast<int>(r[i]);
      cells *= res[i];
      unit[i] = 1.0f / static_cast<float>(res[i]);
    }
  }
  void cell(int c, int out[3]) const {
    int cnt = cells, rem = c;
    for (int i = 0; i < dims; ++i) {
      cnt /= res[i];
      out[i] = rem / cnt;
      rem -= out[i] * cnt;
    }
  }
  float sample_reuse(float* s) const {
    float pdf;
    const int idx = pick.sample_reuse(s[dims - 1], &pdf);
    int c[3];
    cell(idx, c);
    for (int i = 0; i < dims; ++i) s[i] = (s[i] + static_cast<float>(c[i])) * unit[i];
    return pdf;
  }
};

template <class F> struct Film {
  int H, W;
  std::vector<V3<F>> px;
  Film(int h, int w) : H(h), W(w), px(static_cast<size_t>(h) * w, V3<F>(F(0.f))) {}
};

// Film::add (src/camera.h:75-87) for lane results in lane order.
template <class F>
void film_add(Film<F>& film, const std::vector<V3<F>>& value, const std::vector<int>& pix, const std::vector<char>& ok,
              uint32_t spp) {
  std::vector<V3<F>> acc(film.px.size(), V3<F>(F(0.f)));
  for (size_t i = 0; i < value.size(); ++i)
    if (ok[i]) acc[pix[i]] = acc[pix[i]] + value[i];
  for (size_t p = 0; p < acc.size(); ++p) {
    V3<F> v = acc[p];
    if (spp > 1) v = v / F(static_cast<float>(spp));
    film.px[p] = film.px[p] + v;
  }
}

// One lane of render_interior (src/path_integrator.cc:359-368): camera ray
// from the lane's first two draws, hit, trace2light, times the jacobian.
template <class F>
V3<F> lane_radiance(const Scene<F>& sc, const Camera& cam, const oracle_config& cfg, int64_t lane, int* pix,
                    bool* valid) {
  Pcg32 rng = lane_rng(static_cast<uint32_t>(lane));
  const float u0 = rng.next_float();
  const float u1 = rng.next_float();
  V3<float> o, d;
  V3<Dual> o_ad;
  int x, y;
  cam.ray(static_cast<uint32_t>(lane), cfg.spp_interior, u0, u1, &o, &d, &x, &y, &o_ad);
  const Its<F> its = sc.hit(origin_as<F>(o, o_ad), lift<F>(d), true);
  V3<F> L = trace2light<F>(sc, rng, its, its.valid, cfg.mis_light_samples, cfg.mis_bsdf_samples, cfg.max_bounce,
                           cfg.enable_light_visable != 0);
  *pix = y * cam.W + x;
  *valid = its.valid;
  return L * its.jac;
}

// render_interior (src/path_integrator.cc:355-372)
template <class F>
void render_interior(const Scene<F>& sc, const Camera& cam, const oracle_config& cfg, Film<F>& film,
                     float* lane_L) {
  const uint32_t spp = cfg.spp_interior;
  const int64_t N = static_cast<int64_t>(spp) * cam.H * cam.W;
  std::vector<V3<F>> vals(N);
  std::vector<int> pix(N);
  std::vector<char> ok(N);
  parallel_for(N, [&](int64_t b, int64_t e) {
    for (int64_t lane = b; lane < e; ++lane) {
      int p;
      bool valid;
      const V3<F> L = lane_radiance<F>(sc, cam, cfg, lane, &p, &valid);
      vals[lane] = L;
      pix[lane] = p;
      ok[lane] = valid;
      if (lane_L) {
        const V3<float> v = detach(L);
        lane_L[3 * lane] = valid ? v.x : 0.f;
        lane_L[3 * lane + 1] = valid ? v.y : 0.f;
        lane_L[3 * lane + 2] = valid ? v.z : 0.f;
      }
    }
  });
  film_add(film, vals, pix, ok, spp);
}

// preprocess_primary_boundary (src/path_integrator.cc:253-272)
HyperCube preprocess_primary(const Scene<float>& scv, const Camera& cam, const oracle_config& cfg) {
  HyperCube hc;
  hc.init(1, &cfg.primary_edge_hypercube_resolution);
  std::vector<float> rst(hc.cells, 0.f);
  std::vector<Pcg32> rng(hc.cells);
  for (int i = 0; i < hc.cells; ++i) rng[i] = lane_rng(static_cast<uint32_t>(i));
  for (uint32_t r = 0; r < cfg.primary_edge_preprocess_rounds; ++r) {
    parallel_for(hc.cells, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        const float s = (static_cast<float>(i) + rng[i].next_float()) * hc.unit[0];
        BoundaryOut bo;
        const V3<float> v =
            eval_primary<float>(scv, scv, cam, rng[i], s, cfg.mis_light_samples, cfg.max_bounce, false, &bo);
        if (bo.valid) rst[i] = rst[i] + norm(v);
      }
    });
  }
  hc.pick.build(rst.data(), hc.cells);
  return hc;
}

// preprocess_direct_boundary (src/path_integrator.cc:331-353)
HyperCube preprocess_direct(const Scene<float>& scv, const Camera& cam, const oracle_config& cfg) {
  HyperCube hc;
  hc.init(3, cfg.secondary_edge_hypercube_resolution);
  std::vector<float> rst(hc.cells, 0.f);
  std::vector<Pcg32> rng(hc.cells);
  for (int i = 0; i < hc.cells; ++i) rng[i] = lane_rng(static_cast<uint32_t>(i));
  for (uint32_t r = 0; r < cfg.secondary_edge_preprocess_rounds; ++r) {
    parallel_for(hc.cells, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        int c[3];
        hc.cell(static_cast<int>(i), c);
        float s[3];
        for (int k = 0; k < 3; ++k) s[k] = rng[i].next_float();
        for (int k = 0; k < 3; ++k) s[k] = (static_cast<float>(c[k]) + s[k]) * hc.unit[k];
        BoundaryOut bo;
        V3<float> v = eval_direct<float>(scv, cam, s[0], s[1], s[2], false, &bo);
        for (int k = 0; k < 3; ++k)
          if (v[k] < REAL_EPS) v[k] = REAL_EPS;
        if (bo.valid) rst[i] = rst[i] + norm(v);
      }
    });
  }
  hc.pick.build(rst.data(), hc.cells);
  return hc;
}

// render_primary_boundary (src/path_integrator.cc:188-200)
template <class F>
void render_primary(const Scene<F>& sc, const Scene<float>& scv, const Camera& cam, const oracle_config& cfg,
                    const HyperCube& hc, Film<F>& film) {
  const uint32_t spp = cfg.spp_primary_edge;
  const int64_t N = static_cast<int64_t>(spp) * cam.H * cam.W;
  std::vector<V3<F>> vals(N);
  std::vector<int> pix(N);
  std::vector<char> ok(N);
  parallel_for(N, [&](int64_t b, int64_t e) {
    for (int64_t lane = b; lane < e; ++lane) {
      Pcg32 rng = lane_rng(static_cast<uint32_t>(lane));
      float s = rng.next_float();
      const float pdf = hc.sample_reuse(&s);
      BoundaryOut bo;
      V3<F> v = eval_primary<F>(sc, scv, cam, rng, s, cfg.mis_light_samples, cfg.max_bounce, true, &bo);
      bool valid = bo.valid && std::isfinite(val(squared_norm(v)));
      vals[lane] = v / F(pdf);
      pix[lane] = valid ? bo.py * cam.W + bo.px : 0;
      ok[lane] = valid;
    }
  });
  film_add(film, vals, pix, ok, spp);
}

// render_direct_boundary (src/path_integrator.cc:274-285)
template <class F>
void render_direct(const Scene<F>& sc, const Camera& cam, const oracle_config& cfg, const HyperCube& hc,
                   Film<F>& film) {
  const uint32_t spp = cfg.spp_secondary_edge;
  const int64_t N = static_cast<int64_t>(spp) * cam.H * cam.W;
  std::vector<V3<F>> vals(N);
  std::vector<int> pix(N);
  std::vector<char> ok(N);
  parallel_for(N, [&](int64_t b, int64_t e) {
    for (int64_t lane = b; lane < e; ++lane) {
      Pcg32 rng = lane_rng(static_cast<uint32_t>(lane));
      float s[3];
      for (int k = 0; k < 3; ++k) s[k] = rng.next_float();
      const float pdf = hc.sample_reuse(s);
      BoundaryOut bo;
      const V3<F> v = eval_direct<F>(sc, cam, s[0], s[1], s[2], true, &bo);
      vals[lane] = v / F(pdf);
      pix[lane] = bo.valid ? bo.py * cam.W + bo.px : 0;
      ok[lane] = bo.valid;
    }
  });
  film_add(film, vals, pix, ok, spp);
}

Camera make_camera(const oracle_camera* c) {
  Camera cam;
  cam.build(c->look_from, c->look_at, c->up, c->vfov, c->height, c->width);
  if (c->kind == 1) cam.build_ortho(c->view_width);
  cam.build_tangent(c->look_from, c->look_at, c->up, c->kind == 1 ? c->view_width : c->vfov, c->d_params);
  return cam;
}

// ---------------------------------------------------------------------------
// LBVH restatement: 30-bit Morton codes of triangle box centres with bit 30 set
// for triangles whose largest padded extent is <= 1/8 of the centre-bounds extent,
// stable sort,
// Karras (2012) hierarchy over keys (morton << 32 | sorted position).
// ---------------------------------------------------------------------------
inline uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
inline int delta(const std::vector<uint32_t>& m, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = (static_cast<uint64_t>(m[i]) << 32) | static_cast<uint32_t>(i);
  const uint64_t b = (static_cast<uint64_t>(m[j]) << 32) | static_cast<uint32_t>(j);
  return __builtin_clzll(a ^ b);
}

}  // namespace

extern "C" {

int oracle_version() { return 1; }
void oracle_set_threads(int n) { g_threads = n; }

void oracle_pcg_kat(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
  Pcg32 r;
  r.seed(initstate, initseq);
  for (int i = 0; i < n; ++i) out[i] = r.next_u32();
}

void oracle_lane_draws(uint32_t lane0, int n_lanes, int n_draws, uint32_t* out) {
  for (int l = 0; l < n_lanes; ++l) {
    Pcg32 r = lane_rng(lane0 + static_cast<uint32_t>(l));
    for (int k = 0; k < n_draws; ++k) out[static_cast<int64_t>(l) * n_draws + k] = r.next_u32();
  }
}

void oracle_lane_seed(uint32_t lane0, int n_lanes, uint64_t* state, uint64_t* inc) {
  for (int l = 0; l < n_lanes; ++l) {
    Pcg32 r = lane_rng(lane0 + static_cast<uint32_t>(l));
    state[l] = r.state;
    inc[l] = r.inc;
  }
}

void oracle_sincos(const float* x, int n, float* s, float* c) {
  for (int i = 0; i < n; ++i) psdr_sincos(x[i], &s[i], &c[i]);
}

void oracle_camera_matrices(const oracle_camera* c, float* c2w, float* w2c, float* c2r, float* r2c, float* pos) {
  const Camera cam = make_camera(c);
  std::memcpy(c2w, cam.c2w.m, 64);
  std::memcpy(w2c, cam.w2c.m, 64);
  std::memcpy(c2r, cam.c2r.m, 64);
  std::memcpy(r2c, cam.r2c.m, 64);
  pos[0] = cam.pos.x;
  pos[1] = cam.pos.y;
  pos[2] = cam.pos.z;
}

// Camera rays of every lane of Sampler(spp*H*W): the first two draws of each
// lane; do (optional): the origins' tangents along c->d_params.
void oracle_camera_rays(const oracle_camera* c, uint32_t spp, float* o, float* d, int* pix, float* dorig) {
  const Camera cam = make_camera(c);
  const int64_t N = static_cast<int64_t>(spp) * cam.H * cam.W;
  for (int64_t lane = 0; lane < N; ++lane) {
    Pcg32 rng = lane_rng(static_cast<uint32_t>(lane));
    const float u0 = rng.next_float(), u1 = rng.next_float();
    V3<float> oo, dd;
    V3<Dual> oad;
    int x, y;
    cam.ray(static_cast<uint32_t>(lane), spp, u0, u1, &oo, &dd, &x, &y, &oad);
    for (int a = 0; a < 3; ++a) {
      o[3 * lane + a] = oo[a];
      d[3 * lane + a] = dd[a];
      if (dorig) dorig[3 * lane + a] = oad[a].d;
    }
    pix[lane] = y * cam.W + x;
  }
}

// Rays of gen_camera_ray_with_point for n points and the tangents of their
// origins along c->d_params and the points' tangents dpts.
void oracle_point_rays(const oracle_camera* c, const float* pts, const float* dpts, int n, float* o, float* d,
                       float* dorig) {
  const Camera cam = make_camera(c);
  for (int i = 0; i < n; ++i) {
    V3<float> oo, dd;
    V3<Dual> oad;
    cam.point_ray(V3<float>{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]},
                  V3<float>{dpts[3 * i], dpts[3 * i + 1], dpts[3 * i + 2]}, &oo, &dd, &oad);
    for (int a = 0; a < 3; ++a) {
      o[3 * i + a] = oo[a];
      d[3 * i + a] = dd[a];
      dorig[3 * i + a] = oad[a].d;
    }
  }
}

// Interior radiance of the n lanes listed in `lanes` (0 where the camera ray
// misses), for parity checks of renders too large for a full oracle image.
int oracle_render_lanes(const oracle_scene* s, const oracle_config* cfg, const oracle_camera* c,
                        const int64_t* lanes, int64_t n, float* out) {
  const Camera cam = make_camera(c);
  Scene<float> scv;
  scv.build(*s, false);
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      int p;
      bool valid;
      const V3<float> L = lane_radiance<float>(scv, cam, *cfg, lanes[i], &p, &valid);
      for (int a = 0; a < 3; ++a) out[3 * i + a] = valid ? L[a] : 0.f;
    }
  });
  return 0;
}

// SH9 radiance of n unit directions for the 27 coefficients (9 RGB rows).
void oracle_sh9_eval(const float* coeff, const float* dirs, int n, float* out) {
  for (int i = 0; i < n; ++i) {
    const V3<float> r = sh9_radiance<float>(coeff, V3<float>{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]});
    for (int a = 0; a < 3; ++a) out[3 * i + a] = r[a];
  }
}

// Tangents of camera2world, world2camera and raster2camera along c->d_params.
void oracle_camera_tangents(const oracle_camera* c, float* dC, float* dW, float* dR) {
  const Camera cam = make_camera(c);
  std::memcpy(dC, cam.dC.m, 64);
  std::memcpy(dW, cam.dW.m, 64);
  std::memcpy(dR, cam.dR.m, 64);
}

void oracle_project_points(const oracle_camera* c, const float* pts, int n, int* pix_xy, int* valid) {
  const Camera cam = make_camera(c);
  for (int i = 0; i < n; ++i) {
    int x, y;
    valid[i] = cam.ray_with_point(V3<float>{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, &x, &y);
    pix_xy[2 * i] = x;
    pix_xy[2 * i + 1] = y;
  }
}

// Per-triangle derived data (normal, tangent, area) and the edge list.
// tri_out: [n_tris*7] = n(3), t(3), area.  edge_topo: [cap*6] = i, j, l, r, kl, kr;
// edge_geom: [cap*16] = p0, p1, out, nl, nr, len.  Returns the edge count.
int oracle_prepare(const oracle_scene* s, float* tri_out, int* edge_topo, float* edge_geom, int cap_edges) {
  Scene<float> sc;
  sc.build(*s, false);
  for (int f = 0; f < s->n_tris; ++f) {
    const Tri<float>& T = sc.tris[f];
    for (int a = 0; a < 3; ++a) {
      tri_out[7 * f + a] = T.n[a];
      tri_out[7 * f + 3 + a] = T.t[a];
    }
    tri_out[7 * f + 6] = T.area;
  }
  std::vector<EdgeTopo> topo = edge_topology(s->faces, sc.n_edge_tris);
  const int ne = static_cast<int>(topo.size());
  for (int e = 0; e < ne && e < cap_edges; ++e) {
    const EdgeTopo& t = topo[e];
    int* o = edge_topo + 6 * e;
    o[0] = t.i;
    o[1] = t.j;
    o[2] = t.l;
    o[3] = t.r;
    o[4] = t.kl;
    o[5] = t.kr;
    const Edge<float>& E = sc.edges[e];
    float* g = edge_geom + 16 * e;
    for (int a = 0; a < 3; ++a) {
      g[a] = E.p0[a];
      g[3 + a] = E.p1[a];
      g[6 + a] = E.out[a];
      g[9 + a] = E.nl[a];
      g[12 + a] = E.nr[a];
    }
    g[15] = E.len;
  }
  return ne;
}

// Light / edge sampler tables (pmf, cmf) and light power, for parity tests.
void oracle_samplers(const oracle_scene* s, float* light_pmf, float* light_cmf, float* edge_pmf, float* edge_cmf,
                     float* light_power, float* light_area) {
  Scene<float> sc;
  sc.build(*s, false);
  for (int o = 0; o < s->n_objects; ++o) {
    light_pmf[o] = sc.light_sampler.pmf[o];
    light_cmf[o] = sc.light_sampler.cmf[o];
  }
  for (size_t e = 0; e < sc.edges.size(); ++e) {
    edge_pmf[e] = sc.edge_sampler.pmf[e];
    edge_cmf[e] = sc.edge_sampler.cmf[e];
  }
  for (int l = 0; l < s->n_lights; ++l) {
    light_power[l] = sc.lights[l].power;
    light_area[l] = sc.lights[l].total_area;
  }
}

// Closest hit of rays against a triangle soup (global ids = array order).
void oracle_closest_hit(const float* tri9, int n_tris, const float* ro, const float* rd, int n_rays, int brute,
                        int* tri, float* t, float* u, float* v) {
  std::vector<V3<float>> P0(n_tris), P1(n_tris), P2(n_tris);
  for (int f = 0; f < n_tris; ++f) {
    const float* q = tri9 + 9 * f;
    P0[f] = {q[0], q[1], q[2]};
    P1[f] = {q[3], q[4], q[5]};
    P2[f] = {q[6], q[7], q[8]};
  }
  CpuBvh bvh;
  bvh.build(P0, P1, P2);
  parallel_for(n_rays, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const V3<float> o{ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
      const V3<float> d{rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
      const Hit h = brute ? bvh.brute(o, d) : bvh.closest(o, d);
      tri[i] = h.tri;
      t[i] = h.tri >= 0 ? h.t : -1.f;
      u[i] = h.u;
      v[i] = h.v;
    }
  });
}

// Per-ray class of the closest hit: bit 0 = near tie (closest and
// second-closest triangle distances within 4 ulp), bit 1 = grazing (hit
// barycentrics within 1e-6 of an edge).
void oracle_hit_classes(const float* tri9, int n_tris, const float* ro, const float* rd, int n_rays, int* cls) {
  std::vector<V3<float>> P0(n_tris), P1(n_tris), P2(n_tris);
  for (int f = 0; f < n_tris; ++f) {
    const float* q = tri9 + 9 * f;
    P0[f] = {q[0], q[1], q[2]};
    P1[f] = {q[3], q[4], q[5]};
    P2[f] = {q[6], q[7], q[8]};
  }
  CpuBvh bvh;
  bvh.build(P0, P1, P2);
  parallel_for(n_rays, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const V3<float> o{ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
      const V3<float> d{rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
      float t2;
      const Hit h = bvh.closest2(o, d, &t2);
      int c = 0;
      if (h.tri >= 0) {
        int32_t a, b2;
        std::memcpy(&a, &h.t, 4);
        std::memcpy(&b2, &t2, 4);
        if (std::isfinite(t2) && b2 - a <= 4) c |= 1;
        if (std::fmin(std::fmin(h.u, h.v), 1.0f - h.u - h.v) < 1e-6f) c |= 2;
      }
      cls[i] = c;
    }
  });
}

// LBVH over a triangle soup.  Outputs: morton[n] (by triangle), order[n] (sorted
// position -> triangle), child[2*(n-1)] (internal node i: left, right; a leaf at
// sorted position k is encoded as k | 0x80000000), parent[2n-1] (internal nodes
// first, then leaves; root's parent = -1), box[(n-1)*12] (per internal node: left
// child lo3 hi3, right child lo3 hi3; boxes padded by (|lo|+|hi|)*1e-6 + 1e-30).
int oracle_lbvh(const float* tri9, int n, uint32_t* morton, int* order, int* child, int* parent, float* box) {
  if (n < 2) return -1;
  std::vector<float> lo(3 * n), hi(3 * n), cen(3 * n);
  float cmin[3], cmax[3];
  for (int a = 0; a < 3; ++a) {
    cmin[a] = std::numeric_limits<float>::infinity();
    cmax[a] = -std::numeric_limits<float>::infinity();
  }
  for (int f = 0; f < n; ++f)
    for (int a = 0; a < 3; ++a) {
      const float x0 = tri9[9 * f + a], x1 = tri9[9 * f + 3 + a], x2 = tri9[9 * f + 6 + a];
      const float l = std::fmin(x0, std::fmin(x1, x2)), h = std::fmax(x0, std::fmax(x1, x2));
      const float pad = (std::fabs(l) + std::fabs(h)) * 1e-6f + 1e-30f;
      lo[3 * f + a] = l - pad;
      hi[3 * f + a] = h + pad;
      cen[3 * f + a] = (l + h) * 0.5f;
      cmin[a] = std::fmin(cmin[a], cen[3 * f + a]);
      cmax[a] = std::fmax(cmax[a], cen[3 * f + a]);
    }
  for (int f = 0; f < n; ++f) {
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      const float ext = cmax[a] - cmin[a];
      const float x = ext > 0.f ? (cen[3 * f + a] - cmin[a]) / ext : 0.5f;
      q[a] = static_cast<uint32_t>(std::fmin(std::fmax(x * 1024.0f, 0.0f), 1023.0f));
    }
    uint32_t m = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    float cmx = 0.f, emx = 0.f;
    for (int a = 0; a < 3; ++a) {
      emx = std::fmax(emx, cmax[a] - cmin[a]);
      cmx = std::fmax(cmx, hi[3 * f + a] - lo[3 * f + a]);
    }
    if (!(cmx > 0.125f * emx)) m |= 1u << 30;  // small triangles sort after the large ones
    morton[f] = m;
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return morton[a] < morton[b]; });
  std::vector<uint32_t> ms(n);
  for (int k = 0; k < n; ++k) {
    order[k] = ord[k];
    ms[k] = morton[ord[k]];
  }
  for (int k = 0; k < 2 * n - 1; ++k) parent[k] = -1;
  for (int i = 0; i < n - 1; ++i) {
    const int d = (delta(ms, n, i, i + 1) - delta(ms, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(ms, n, i, i - d);
    int lmax = 2;
    while (delta(ms, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (delta(ms, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(ms, n, i, j);
    int s = 0, div = 2, t;
    do {
      t = (l + div - 1) / div;
      if (delta(ms, n, i, i + (s + t) * d) > dnode) s += t;
      div *= 2;
    } while (t > 1);
    const int g = i + s * d + std::min(d, 0);
    const int left = (std::min(i, j) == g) ? (g | static_cast<int>(0x80000000u)) : g;
    const int right = (std::max(i, j) == g + 1) ? ((g + 1) | static_cast<int>(0x80000000u)) : g + 1;
    child[2 * i] = left;
    child[2 * i + 1] = right;
    parent[(left < 0) ? (n - 1 + (left & 0x7fffffff)) : left] = i;
    parent[(right < 0) ? (n - 1 + (right & 0x7fffffff)) : right] = i;
  }
  // boxes, bottom-up by recursion from the root
  std::function<void(int, float*, float*)> refit = [&](int c, float* bl, float* bh) {
    if (c < 0) {
      const int f = ord[c & 0x7fffffff];
      for (int a = 0; a < 3; ++a) {
        bl[a] = lo[3 * f + a];
        bh[a] = hi[3 * f + a];
      }
      return;
    }
    float l0[3], h0[3], l1[3], h1[3];
    refit(child[2 * c], l0, h0);
    refit(child[2 * c + 1], l1, h1);
    float* b = box + 12 * c;
    for (int a = 0; a < 3; ++a) {
      b[a] = l0[a];
      b[3 + a] = h0[a];
      b[6 + a] = l1[a];
      b[9 + a] = h1[a];
      bl[a] = std::fmin(l0[a], l1[a]);
      bh[a] = std::fmax(h0[a], h1[a]);
    }
  };
  float rl[3], rh[3];
  refit(0, rl, rh);
  return 0;
}

// Render one camera.  mode 0: renderC (value only).  mode 1: renderD with the
// forward-mode tangent given in the scene's d_* arrays; the tangent image is
// split into interior / primary-boundary / direct-boundary parts (any may be
// NULL to skip that estimator).  lane_L (optional): per-lane interior radiance.
void oracle_set_fd_measure(int on) { g_fd_measure = on != 0; }
void oracle_set_direct_unit_normal(int on) { g_direct_unit = on != 0; }

// DiffuseBsdf value for n cases of (color[3], sigma, N[3], V[3], L[3]) -> out[3].
void oracle_diffuse_bsdf(const float* cases, int n, float* out) {
  for (int i = 0; i < n; ++i) {
    const float* c = cases + 13 * i;
    const V3<float> v = diffuse_bsdf_value(V3<float>(c[0], c[1], c[2]), c[3], V3<float>(c[4], c[5], c[6]),
                                           V3<float>(c[7], c[8], c[9]), V3<float>(c[10], c[11], c[12]));
    out[3 * i] = v.x;
    out[3 * i + 1] = v.y;
    out[3 * i + 2] = v.z;
  }
}

int oracle_render(const oracle_scene* s, const oracle_config* cfg, const oracle_camera* c, int mode, float* img,
                  float* d_int, float* d_pri, float* d_dir, float* lane_L) {
  const Camera cam = make_camera(c);
  const int P = cam.H * cam.W;
  Scene<float> scv;
  scv.build(*s, false);
  if (mode == 0) {
    Film<float> film(cam.H, cam.W);
    for (uint32_t n = 0; n < cfg->n_pass; ++n) render_interior<float>(scv, cam, *cfg, film, n == 0 ? lane_L : nullptr);
    for (int p = 0; p < P; ++p) {
      V3<float> v = film.px[p];
      if (cfg->n_pass > 1) v = v / static_cast<float>(cfg->n_pass);
      for (int a = 0; a < 3; ++a) img[3 * p + a] = v[a];
    }
    return 0;
  }
  Scene<Dual> sc;
  sc.build(*s, true);
  const bool boundary = !sc.edges.empty();
  HyperCube h1, h3;
  if (boundary && d_pri) h1 = preprocess_primary(scv, cam, *cfg);
  if (boundary && d_dir) h3 = preprocess_direct(scv, cam, *cfg);
  Film<Dual> fi(cam.H, cam.W), fp(cam.H, cam.W), fd(cam.H, cam.W);
  for (uint32_t n = 0; n < cfg->n_pass; ++n) {
    render_interior<Dual>(sc, cam, *cfg, fi, n == 0 ? lane_L : nullptr);
    if (boundary && d_pri) render_primary<Dual>(sc, scv, cam, *cfg, h1, fp);
    if (boundary && d_dir) render_direct<Dual>(sc, cam, *cfg, h3, fd);
  }
  const float np = static_cast<float>(cfg->n_pass);
  for (int p = 0; p < P; ++p)
    for (int a = 0; a < 3; ++a) {
      Dual vi = fi.px[p][a], vp = fp.px[p][a], vd = fd.px[p][a];
      Dual tot = (vi + vp) + vd;
      if (cfg->n_pass > 1) {
        tot = tot / Dual(np);
        vi = vi / Dual(np);
        vp = vp / Dual(np);
        vd = vd / Dual(np);
      }
      img[3 * p + a] = tot.v;
      if (d_int) d_int[3 * p + a] = vi.d;
      if (d_pri) d_pri[3 * p + a] = vp.d;
      if (d_dir) d_dir[3 * p + a] = vd.d;
    }
  return 0;
}

}  // extern "C"
