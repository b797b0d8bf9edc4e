// Path-replay adjoint of the interior estimator (trace2light + render_interior,
// src/path_integrator.cc:80-186, 355-372) for renderD's backward pass.
//
// One thread per sample lane.  Inputs: dLoss/dImage, the per-lane hit cache of
// the forward render (no ray is traced again), the scene.  A forward sweep
// replays the lane's PCG32 stream and records, per depth, the PCG state and the
// throughput T_d; a reverse sweep recomputes each depth's contribution c_d and
// throughput factor t_d (L = sum_d T_d c_d, T_{d+1} = T_d t_d) and pushes
//     adj(c_d) = g T_d,   adj(t_d) = A_{d+1} T_d,   A_d = g c_d + A_{d+1} t_d
// through hand-written VJPs of the reference's expressions.  Exactly the
// reference's detach points are honoured: pdfs and MIS weights are constants,
// hit barycentrics are constants (points move with the surface), emission,
// albedo / roughness texels, normals, tangents, points, squared distances and
// jacobians (A / detach(A)) are differentiable.
//
// Gradients: vertices and UV-texels by global fp32 atomics; light emission and
// constant textures (one address hit by every lane) by LDS accumulation per
// block, flushed with one global atomic per entry.
#include <hip/hip_runtime.h>

#include <cstring>

#include "cam_grad.h"
#include "grad_cache.h"
#include "render.h"
#include "shading.h"

namespace psdr {

namespace {

constexpr int ADJ_THREADS = 256;
constexpr int HOT_MAX = 1024;  // LDS accumulators for emission + constant textures
// LDS aggregation of vertex (array 0) and RGB-texel (array 1) gradients of one
// workgroup's contiguous lane range (only k_adjoint uses these arrays)
constexpr int ADJ_SLOTS = 1024;
__shared__ int adj_keys[ADJ_SLOTS];
__shared__ float adj_vals[3 * ADJ_SLOTS];
__device__ __forceinline__ LdsGradCache<ADJ_SLOTS> adj_cache() { return LdsGradCache<ADJ_SLOTS>{adj_keys, adj_vals}; }

struct AdjParams {
  SceneView sv;
  CameraK cam;
  int D;
  uint32_t nL, nB;
  int elv;
  uint32_t spp;
  int Kmax;
  int64_t N;
  const int* ctri;
  const float* cu;
  const float* cv;
  const float* cT;      // [D][3][N] throughput per depth (forward render)
  const uint64_t* crs;  // [D][N] PCG state per depth
  const float* cL;      // [3][N] path radiance
  const int* cinfo;     // [N] depths | 256 light camera vertex | 512 radiance not finite
  const float* dLdI;
  float* grad_v;
  float* grad_t;
  float* grad_e;
  float* cam_acc;  // camera matrix adjoints (CAM_ACC floats), or null
  int hot_n;  // 3*n_lights + 4*n_objects if it fits in HOT_MAX, else 0
};

__device__ __forceinline__ f3 vjp_normalize(f3 x, f3 gy) {
  const float r = 1.0f / sqrtf(sqnorm(x));
  const f3 y = x * r;
  return (gy - y * dot(y, gy)) * r;
}

__device__ __forceinline__ void acc_hot(float* hot, int hot_n, float* global, int slot, int gidx, float v) {
  if (v == 0.f) return;
  if (hot_n > 0) atomicAdd(hot + slot, v);
  else atomicAdd(global + gidx, v);
}

struct VertexAdj {
  f3 p, n, t;  // adjoints of the point, shading normal, tangent (after the front flip)
  float area;  // adjoint of the triangle area (through the jacobian A / detach(A))
};
__device__ __forceinline__ VertexAdj zero_adj() { return VertexAdj{splat(0.f), splat(0.f), splat(0.f), 0.f}; }
__device__ __forceinline__ bool nonzero(const VertexAdj& a) {
  return (a.p.x != 0.f || a.p.y != 0.f || a.p.z != 0.f) || (a.n.x != 0.f || a.n.y != 0.f || a.n.z != 0.f) ||
         (a.t.x != 0.f || a.t.y != 0.f || a.t.z != 0.f) || a.area != 0.f;
}

// VJP of a triangle record (src/mesh.cc:38-67) and of the barycentric point
// p = (P0 + u (P1-P0)) + v (P2-P0) (src/scene.cc:130-133) into vertex gradients.
// `flip` = -1 when the record's normal/tangent were negated to face the ray.
__device__ void tri_vjp(const AdjParams& P, int tri, float u, float v, float flip, const VertexAdj& g) {
  const SceneView& sv = P.sv;
  const int obj = sv.tri_obj[tri];
  if (!(sv.objs[obj].grad_mask & PSDR_GRAD_VERTICES)) return;
  const float4 g0 = sv.tri_geom[5 * tri], g1 = sv.tri_geom[5 * tri + 1], g2 = sv.tri_geom[5 * tri + 2];
  const f3 P0 = xyz4(g0), P1 = xyz4(g1), P2 = xyz4(g2);
  f3 gP0 = g.p * ((1.0f - u) - v);
  f3 gP1 = g.p * u;
  f3 gP2 = g.p * v;
  const f3 a = P2 - P1, b = P0 - P2, c = P1 - P0;
  f3 ga = splat(0.f), gb = splat(0.f), gc = splat(0.f);
  if (g.n.x != 0.f || g.n.y != 0.f || g.n.z != 0.f) {
    const f3 gr = vjp_normalize(cross(c, b), g.n * flip);
    gc = gc + cross(b, gr);
    gb = gb + cross(gr, c);
  }
  if (g.t.x != 0.f || g.t.y != 0.f || g.t.z != 0.f) {
    const float* uv = sv.tri_uv + 6 * tri;
    const float duv1y = uv[3] - uv[1], duv2y = uv[5] - uv[1];
    const f3 gr = vjp_normalize(duv2y * c + duv1y * b, g.t * flip);
    gc = gc + duv2y * gr;
    gb = gb + duv1y * gr;
  }
  if (g.area != 0.f) {
    const float la = norm(a), lb = norm(b), lc = norm(c);
    const float p = ((la + lb) + lc) / 2.0f;
    const float q1 = p - la, q2 = p - lb, q3 = p - lc;
    const float Q = ((p * q1) * q2) * q3;
    if (Q > 0.f) {
      const float dAdQ = 0.5f / sqrtf(Q);
      const float s = g.area * dAdQ * 0.5f;
      const float qqq = q1 * q2 * q3;
      const float dla = qqq - p * q2 * q3 + p * q1 * q3 + p * q1 * q2;
      const float dlb = qqq + p * q2 * q3 - p * q1 * q3 + p * q1 * q2;
      const float dlc = qqq + p * q2 * q3 + p * q1 * q3 - p * q1 * q2;
      ga = ga + a * (s * dla / la);
      gb = gb + b * (s * dlb / lb);
      gc = gc + c * (s * dlc / lc);
    }
  }
  gP2 = gP2 + ga;
  gP1 = gP1 - ga;
  gP0 = gP0 + gb;
  gP2 = gP2 - gb;
  gP1 = gP1 + gc;
  gP0 = gP0 - gc;
  const int* f = sv.faces + 3 * tri;
  const f3 gs[3] = {gP0, gP1, gP2};
  LdsGradCache<ADJ_SLOTS> cache = adj_cache();
#pragma unroll
  for (int k = 0; k < 3; ++k) cache.add(P.grad_v, 0, 3 * static_cast<int64_t>(f[k]), gs[k]);
}

// flip sign of a reconstructed intersection: its.n = flip * record normal
__device__ __forceinline__ float flip_of(const SceneView& sv, const Its& it) {
  const float4 g3 = sv.tri_geom[5 * it.tri + 3];
  return (it.n.x == g3.x && it.n.y == g3.y && it.n.z == g3.z) ? 1.f : -1.f;
}

// VJP of Material::eval at (its, L) w.r.t. texels, the shading normal, L and its.wi.
struct BsdfAdj {
  f3 N, L, V;
};
__device__ BsdfAdj bsdf_vjp(const AdjParams& P, float* hot, const Its& it, f3 L, f3 g) {
  const SceneView& sv = P.sv;
  const ObjectK& ob = sv.objs[it.obj];
  BsdfAdj r{splat(0.f), splat(0.f), splat(0.f)};
  const int co = tex3_offset(ob.alb_off, ob.alb_h, ob.alb_w, it.uv);
  f3 gcol;
  if (ob.mat == PSDR_MAT_DIFFUSE_BSDF) {
    const f3 color = tex3(sv.texels, co);
    const int so = tex1_offset(ob.rough_off, ob.rough_h, ob.rough_w, it.uv);
    const float sigma = sv.texels[so];
    const float kk = (3.0f * REAL_PI - 4.0f) / 6.0f;
    const float div = 1.0f / (REAL_PI + kk * sigma);
    const float a = div, b = sigma * div;
    const f3 N = shading_normal(sv, ob, it);
    const float nl_raw = dot(N, L), nv_raw = dot(N, it.wi);
    const float NdotL = nl_raw < 0.f ? 0.f : nl_raw;
    const float NdotV = nv_raw < 0.f ? 0.f : nv_raw;
    const float VdotL = dot(it.wi, L);
    float t = VdotL - NdotL * NdotV;
    const bool tpos = t > 0.f;
    const bool pickL = NdotL >= NdotV;
    if (tpos) t = (pickL ? NdotL : NdotV) + REAL_EPS;
    gcol = g * (a + b * t);
    const float s = (g.x * color.x + g.y * color.y) + g.z * color.z;
    const float ga = s, gb = s * t, gt = s * b;
    if (ob.grad_mask & PSDR_GRAD_ROUGHNESS) {
      const float dadsig = -kk * div * div;
      const float gsig = ga * dadsig + gb * (div + sigma * dadsig);
      if (ob.rough_h == 0) acc_hot(hot, P.hot_n, P.grad_t, 3 * sv.n_lights + 4 * it.obj + 3, so, gsig);
      else atomicAdd(P.grad_t + so, gsig);
    }
    float gNL = 0.f, gNV = 0.f, gVL = 0.f;
    if (tpos) {
      if (pickL) gNL = gt;
      else gNV = gt;
    } else {
      gVL = gt;
      gNL = -gt * NdotV;
      gNV = -gt * NdotL;
    }
    if (nl_raw > 0.f) {
      r.N = r.N + L * gNL;
      r.L = r.L + N * gNL;
    }
    if (nv_raw > 0.f) {
      r.N = r.N + it.wi * gNV;
      r.V = r.V + N * gNV;
    }
    r.V = r.V + L * gVL;
    r.L = r.L + it.wi * gVL;
  } else {
    gcol = g * REAL_INV_PI;
  }
  if (ob.grad_mask & PSDR_GRAD_ALBEDO) {
    if (ob.alb_h == 0) {
      acc_hot(hot, P.hot_n, P.grad_t, 3 * sv.n_lights + 4 * it.obj + 0, co + 0, gcol.x);
      acc_hot(hot, P.hot_n, P.grad_t, 3 * sv.n_lights + 4 * it.obj + 1, co + 1, gcol.y);
      acc_hot(hot, P.hot_n, P.grad_t, 3 * sv.n_lights + 4 * it.obj + 2, co + 2, gcol.z);
    } else {
      adj_cache().add(P.grad_t, 1, co, gcol);
    }
  }
  return r;
}

// Adjoint gN of the shading normal into the vertex adjoints: the geometric
// normal itself, or through the normal map N = normalize(M), M = (T x + B y) +
// Ng z, B = Ng x T, (x, y, z) = 2 n - 1 into T, Ng and the normal-map texel.
__device__ void shading_normal_vjp(const AdjParams& P, const Its& it, f3 gN, VertexAdj& cur) {
  const SceneView& sv = P.sv;
  const ObjectK& ob = sv.objs[it.obj];
  if (!(ob.mat == PSDR_MAT_DIFFUSE_BSDF && ob.nrm_off >= 0)) {
    cur.n = cur.n + gN;
    return;
  }
  const int o = tex3_offset(ob.nrm_off, ob.nrm_h, ob.nrm_w, it.uv);
  const f3 tn = tex3(sv.texels, o);
  const f3 v = tn * 2.0f - splat(1.0f);
  const f3 B = cross(it.n, it.t);
  const f3 M = (it.t * v.x + B * v.y) + it.n * v.z;
  const f3 gM = vjp_normalize(M, gN);
  const f3 gB = gM * v.y;
  cur.t = cur.t + gM * v.x + cross(gB, it.n);
  cur.n = cur.n + gM * v.z + cross(it.t, gB);
  if (ob.grad_mask & PSDR_GRAD_NORMAL)
    adj_cache().add(P.grad_t, 1, o, f3{dot(gM, it.t), dot(gM, B), dot(gM, it.n)} * 2.0f);
}

// Environment part of radiance_vjp (compiled only into the replay kernel for
// scenes with an environment light, whose register budget it would otherwise
// tighten).
__device__ __forceinline__ f3 env_radiance_vjp(const AdjParams& P, float* hot, int obj, const Its& it, f3 g) {
  const SceneView& sv = P.sv;
  const ObjectK& ob = sv.objs[obj];
  if (ob.mat == PSDR_MAT_ENV_MAP) {
    if (ob.grad_mask & PSDR_GRAD_ALBEDO) {
      const int co = tex3_offset(ob.alb_off, ob.alb_h, ob.alb_w, it.uv);
      if (ob.alb_h == 0) {
        acc_hot(hot, P.hot_n, P.grad_t, 3 * sv.n_lights + 4 * obj + 0, co + 0, g.x);
        acc_hot(hot, P.hot_n, P.grad_t, 3 * sv.n_lights + 4 * obj + 1, co + 1, g.y);
        acc_hot(hot, P.hot_n, P.grad_t, 3 * sv.n_lights + 4 * obj + 2, co + 2, g.z);
      } else {
        adj_cache().add(P.grad_t, 1, co, g);
      }
    }
    return splat(0.f);
  }
  if (ob.mat == PSDR_MAT_ENV_SH9) {
    const Sh9Const c = sh9_const();
    const f3 w = it.wi;
    const float* C = sv.texels + ob.alb_off;
    float sk[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) sk[k] = dot(g, f3{C[3 * k], C[3 * k + 1], C[3 * k + 2]});
    if (ob.grad_mask & PSDR_GRAD_ALBEDO) {
      const Sh9Basis b = sh9_basis(w);
      for (int k = 0; k < 9; ++k) {
        const int o = ob.alb_off + 3 * k;
        atomicAdd(P.grad_t + o, g.x * b.y[k]);
        atomicAdd(P.grad_t + o + 1, g.y * b.y[k]);
        atomicAdd(P.grad_t + o + 2, g.z * b.y[k]);
      }
    }
    // d Y_k / d w of Y = (c0, c1 y, c1 z, c1 x, c2 xy, c2 yz, c3 (2zz - xx - yy), c2 zx, c4 (xx - yy))
    const float gx = ((sk[3] * c.c1 + sk[4] * c.c2 * w.y) + (sk[7] * c.c2 * w.z - sk[6] * c.c3 * 2.0f * w.x)) +
                     sk[8] * c.c4 * 2.0f * w.x;
    const float gy = ((sk[1] * c.c1 + sk[4] * c.c2 * w.x) + (sk[5] * c.c2 * w.z - sk[6] * c.c3 * 2.0f * w.y)) -
                     sk[8] * c.c4 * 2.0f * w.y;
    const float gz = (sk[2] * c.c1 + sk[5] * c.c2 * w.y) + (sk[6] * c.c3 * 4.0f * w.z + sk[7] * c.c2 * w.x);
    return f3{gx, gy, gz};
  }
  return splat(0.f);
}

// VJP of light_radiance(sv, l, it) for the adjoint g into the area light's
// emission, the cube map's texel or the SH9 coefficients; returns the adjoint
// of it.wi (nonzero for SH9 only).
template <bool ENV>
__device__ __forceinline__ f3 radiance_vjp(const AdjParams& P, float* hot, int l, const Its& it, f3 g) {
  const SceneView& sv = P.sv;
  const int obj = sv.lights[l].obj;
  const ObjectK& ob = sv.objs[obj];
  if constexpr (ENV) {
    if (ob.mat == PSDR_MAT_ENV_MAP || ob.mat == PSDR_MAT_ENV_SH9) return env_radiance_vjp(P, hot, obj, it, g);
  }
  if (ob.grad_mask & PSDR_GRAD_EMIT) {
    acc_hot(hot, P.hot_n, P.grad_e, 3 * l, 3 * l, g.x);
    acc_hot(hot, P.hot_n, P.grad_e, 3 * l + 1, 3 * l + 1, g.y);
    acc_hot(hot, P.hot_n, P.grad_e, 3 * l + 2, 3 * l + 2, g.z);
  }
  return splat(0.f);
}

__device__ __forceinline__ bool nonzero3(f3 a) { return a.x != 0.f || a.y != 0.f || a.z != 0.f; }

struct DepthOut {
  f3 c;       // sum of weighted NEE + BSDF-MIS values at this depth
  f3 t;       // extension throughput factor
  bool ext;   // extension succeeded (the path continues)
  int ntri;   // next vertex hit
  float nu, nv;
  VertexAdj cur;   // adjoint of this vertex (point, normal, tangent)
  f3 gO;           // adjoint of this vertex's ray origin (through wi)
  VertexAdj next;  // adjoint of the next vertex (point, normal, area)
};

__device__ __forceinline__ void draw_cached(const AdjParams& P, int64_t slot, int64_t lane, int* t, float* u,
                                            float* v) {
  const int64_t i = slot * P.N + lane;
  *t = P.ctri[i];
  *u = P.cu[i];
  *v = P.cv[i];
}

// One depth of trace2light at vertex `its` (state: rng at the depth's first
// draw).  With grad, adjoints adj_c / adj_t are pushed through this depth.
template <bool ENV>
__device__ __forceinline__ DepthOut depth_step(const AdjParams& P, float* hot, const Its& its, Rng rng, int d,
                                               int64_t lane,
                               bool grad, f3 adj_c, f3 adj_t) {
  const SceneView& sv = P.sv;
  DepthOut out;
  out.c = splat(0.f);
  out.t = splat(0.f);
  out.ext = false;
  out.ntri = -1;
  out.nu = out.nv = 0.f;
  out.cur = zero_adj();
  out.next = zero_adj();
  out.gO = splat(0.f);
  const int64_t cslot = 1 + static_cast<int64_t>(d) * P.Kmax;
  int k = 0;
  f3 gV = splat(0.f);  // adjoint of its.wi
  // ---- next-event estimation
  for (uint32_t i = 0; i < P.nL; ++i, ++k) {
    float ru = rng.next();
    const float rv = rng.next();
    float pdf_choose;
    const int obj = discrete_sample(sv.obj_pmf, sv.obj_cmf, sv.n_objects, &ru, &pdf_choose);
    const int l = sv.objs[obj].light;
    if (l < 0) continue;
    const LightPoint lp = sample_light_point(sv, l, ru, rv);
    const float pdf_light = lp.pdf * pdf_choose;
    const f3 dvec = lp.p - its.p;
    const f3 dir = normalize(dvec);
    int ht;
    float hu, hv;
    draw_cached(P, cslot + k, lane, &ht, &hu, &hv);
    const Its itl = reconstruct(sv, ht, hu, hv, its.p, true);
    if (!(itl.valid && itl.is_light && sqnorm(itl.p - lp.p) < REAL_EPS)) continue;
    const f3 emit = light_radiance(sv, l, itl);
    const f3 bsdf = mat_eval(sv, its, dir);
    const float tmp = dot(itl.n, -dir) / itl.d2;
    const float nd = dot(its.n, dir);
    const float G = nd * tmp;
    const float Jl = itl.area / itl.area;
    const f3 value = (((emit * G) * bsdf) * Jl) / pdf_light;
    float weight = 1.0f / static_cast<float>(P.nL);
    if (P.nB > 0) weight = weight * mis_weight(pdf_light, mat_pdf(its, dir) * tmp, P.nL, P.nB);
    out.c = out.c + value * weight;
    if (!grad) continue;
    const f3 av = adj_c * (weight / pdf_light);
    const f3 gwl = radiance_vjp<ENV>(P, hot, l, itl, av * bsdf * (G * Jl));
    const f3 gbsdf = av * emit * (G * Jl);
    const float gG = dot(av, emit * bsdf) * Jl;
    const float gJ = dot(av, emit * bsdf) * G;
    const BsdfAdj ba = bsdf_vjp(P, hot, its, dir, gbsdf);
    shading_normal_vjp(P, its, ba.N, out.cur);
    out.cur.n = out.cur.n + dir * (gG * tmp);
    gV = gV + ba.V;
    f3 gdir = ba.L + its.n * (gG * tmp);
    const float gtmp = gG * nd;
    VertexAdj lh = zero_adj();
    lh.n = (-dir) * (gtmp / itl.d2);
    gdir = gdir + (-itl.n) * (gtmp / itl.d2);
    const float gd2 = -gtmp * tmp / itl.d2;
    const f3 gpl = (itl.p - its.p) * (2.0f * gd2);
    lh.p = gpl;
    out.cur.p = out.cur.p - gpl;
    lh.area = gJ / itl.area;
    if (nonzero3(gwl)) {  // the light hit's wi = -normalize(itl.p - its.p)
      const f3 gx = vjp_normalize(itl.p - its.p, -gwl);
      lh.p = lh.p + gx;
      out.cur.p = out.cur.p - gx;
    }
    const f3 gdv = vjp_normalize(dvec, gdir);
    out.cur.p = out.cur.p - gdv;
    VertexAdj ls = zero_adj();
    ls.p = gdv;
    tri_vjp(P, itl.tri, hu, hv, flip_of(sv, itl), lh);
    tri_vjp(P, lp.tri, lp.bu, lp.bv, 1.f, ls);
  }
  // ---- BSDF sampling with MIS
  for (uint32_t i = 0; i < P.nB; ++i, ++k) {
    const float ru = rng.next();
    const float rv = rng.next();
    const Sampled smp = mat_sample(sv, its, ru, rv);
    int ht;
    float hu, hv;
    draw_cached(P, cslot + k, lane, &ht, &hu, &hv);
    const Its itl = reconstruct(sv, ht, hu, hv, its.p, true);
    if (!(itl.valid && itl.is_light)) continue;
    const int l = itl.light;
    const f3 emit = light_radiance(sv, l, itl);
    const float tmp = dot(itl.n, -smp.wo) / itl.d2;
    const float nd = dot(its.n, smp.wo);
    const float G = nd * tmp;
    const float pdf_bsdf = smp.pdf * tmp;
    const float Jl = itl.area / itl.area;
    const f3 value = (((emit * G) * smp.bsdf) * Jl) / pdf_bsdf;
    float weight = 1.0f / static_cast<float>(P.nB);
    if (P.nL > 0) {
      const float pdf_light = (sv.light_power[l] / *sv.lights_total) * (1.0f / sv.light_area[l]);
      weight = weight * mis_weight(pdf_bsdf, pdf_light, P.nB, P.nL);
    }
    out.c = out.c + value * weight;
    if (!grad) continue;
    const f3 av = adj_c * (weight / pdf_bsdf);
    const f3 gwl = radiance_vjp<ENV>(P, hot, l, itl, av * smp.bsdf * (G * Jl));
    const f3 gbsdf = av * emit * (G * Jl);
    const float gG = dot(av, emit * smp.bsdf) * Jl;
    const float gJ = dot(av, emit * smp.bsdf) * G;
    const BsdfAdj ba = bsdf_vjp(P, hot, its, smp.wo, gbsdf);
    shading_normal_vjp(P, its, ba.N, out.cur);
    out.cur.n = out.cur.n + smp.wo * (gG * tmp);
    gV = gV + ba.V;
    f3 gwo = ba.L + its.n * (gG * tmp);
    const float gtmp = gG * nd;
    VertexAdj lh = zero_adj();
    lh.n = (-smp.wo) * (gtmp / itl.d2);
    gwo = gwo + (-itl.n) * (gtmp / itl.d2);
    const float gd2 = -gtmp * tmp / itl.d2;
    const f3 gpl = (itl.p - its.p) * (2.0f * gd2);
    lh.p = gpl;
    out.cur.p = out.cur.p - gpl;
    lh.area = gJ / itl.area;
    if (nonzero3(gwl)) {  // the light hit's wi = -normalize(itl.p - its.p)
      const f3 gx = vjp_normalize(itl.p - its.p, -gwl);
      lh.p = lh.p + gx;
      out.cur.p = out.cur.p - gx;
    }
    // wo = normalize(u T + v B + cos N), B = cross(Ng, T), N the shading normal
    const f3 B = cross(its.n, its.t);
    const f3 wu = (smp.u * its.t + smp.v * B) + smp.cos_t * shading_normal(sv, sv.objs[its.obj], its);
    const f3 gwu = vjp_normalize(wu, gwo);
    const f3 gB = gwu * smp.v;
    out.cur.t = out.cur.t + gwu * smp.u + cross(gB, its.n);
    shading_normal_vjp(P, its, gwu * smp.cos_t, out.cur);
    out.cur.n = out.cur.n + cross(its.t, gB);
    tri_vjp(P, itl.tri, hu, hv, flip_of(sv, itl), lh);
  }
  // ---- extension
  {
    const float ru = rng.next();
    const float rv = rng.next();
    if (d < P.D - 1) {
      const Sampled smp = mat_sample(sv, its, ru, rv);
      int ht;
      float hu, hv;
      draw_cached(P, cslot + k, lane, &ht, &hu, &hv);
      const Its itn = reconstruct(sv, ht, hu, hv, its.p, true);
      if (smp.valid && itn.valid) {
        out.ext = true;
        out.ntri = ht;
        out.nu = hu;
        out.nv = hv;
        const float tmp = dot(itn.n, -smp.wo) / itn.d2;
        const float nd = dot(its.n, smp.wo);
        const float G = nd * tmp;
        const float pdf = smp.pdf * tmp;
        const float Jn = itn.area / itn.area;
        out.t = ((smp.bsdf * Jn) * G) / pdf;
        if (grad && nonzero3(adj_t)) {
          const f3 at = adj_t / pdf;
          const f3 gbsdf = at * (Jn * G);
          const float gJ = dot(at, smp.bsdf) * G;
          const float gG = dot(at, smp.bsdf) * Jn;
          const BsdfAdj ba = bsdf_vjp(P, hot, its, smp.wo, gbsdf);
          shading_normal_vjp(P, its, ba.N, out.cur);
          out.cur.n = out.cur.n + smp.wo * (gG * tmp);
          gV = gV + ba.V;
          f3 gwo = ba.L + its.n * (gG * tmp);
          const float gtmp = gG * nd;
          out.next.n = (-smp.wo) * (gtmp / itn.d2);
          gwo = gwo + (-itn.n) * (gtmp / itn.d2);
          const float gd2 = -gtmp * tmp / itn.d2;
          const f3 gpn = (itn.p - its.p) * (2.0f * gd2);
          out.next.p = gpn;
          out.cur.p = out.cur.p - gpn;
          out.next.area = gJ / itn.area;
          const f3 B = cross(its.n, its.t);
          const f3 wu = (smp.u * its.t + smp.v * B) + smp.cos_t * shading_normal(sv, sv.objs[its.obj], its);
          const f3 gwu = vjp_normalize(wu, gwo);
          const f3 gB = gwu * smp.v;
          out.cur.t = out.cur.t + gwu * smp.u + cross(gB, its.n);
          shading_normal_vjp(P, its, gwu * smp.cos_t, out.cur);
          out.cur.n = out.cur.n + cross(its.t, gB);
        }
      }
    }
  }
  if (grad && (gV.x != 0.f || gV.y != 0.f || gV.z != 0.f)) {
    // wi = -normalize(p - O); p - O = -wi * sqrt(d2)
    const f3 wu = its.wi * (-sqrtf(its.d2));
    const f3 gw = vjp_normalize(wu, gV * -1.0f);
    out.cur.p = out.cur.p + gw;
    out.gO = out.gO - gw;
  }
  return out;
}

// Lanes are replayed in chunks of ADJ_CHUNK lanes of the workgroup's range,
// reordered inside the chunk by their number of depths (counting sort into
// ADJ_BINS bins): a wave's loop over depths runs as long as its longest path,
// so waves of paths of one length idle least.  The order within a bin is the
// lanes' own order, so the reordering is deterministic.
constexpr int ADJ_SUB = 4;                          // lanes per thread per chunk
constexpr int ADJ_CHUNK = ADJ_SUB * ADJ_THREADS;
constexpr int ADJ_WAVES = ADJ_THREADS / 64;
constexpr int ADJ_BINS = 8;

template <int DMAX>
__device__ __forceinline__ int adj_depth_key(const AdjParams& P, int64_t lane, int64_t lane_end) {
  if (lane >= lane_end) return 0;
  const int info = P.cinfo[lane];
  if (info & 512) return 0;
  return min(min(info & 255, DMAX), ADJ_BINS - 1);
}

template <int DMAX, bool ENV>
__global__ __launch_bounds__(ADJ_THREADS) void k_adjoint(AdjParams P) {
  __shared__ float hot[HOT_MAX];
  __shared__ float cacc[CAM_ACC];
  __shared__ int order[ADJ_CHUNK];
  __shared__ int bins[ADJ_SUB * ADJ_WAVES * ADJ_BINS];
  for (int i = threadIdx.x; i < P.hot_n; i += blockDim.x) hot[i] = 0.f;
  cam_acc_init(cacc);
  adj_cache().init();
  __syncthreads();
  const SceneView& sv = P.sv;
  // each workgroup replays a contiguous lane range: neighbouring pixels share
  // triangles and texels, so their scatters aggregate in LDS
  const int64_t per = ((P.N + gridDim.x - 1) / gridDim.x + ADJ_THREADS - 1) / ADJ_THREADS * ADJ_THREADS;
  const int64_t lane_end = min(P.N, static_cast<int64_t>(blockIdx.x) * per + per);
  const int wave = threadIdx.x >> 6;
  const uint64_t lt = (1ull << (threadIdx.x & 63)) - 1ull;
  for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * per; c0 < lane_end; c0 += ADJ_CHUNK) {
    // rank of each lane within its bin: lanes of earlier sub-rows, then
    // earlier waves of this sub-row, then earlier lanes of this wave
    int key[ADJ_SUB], rank[ADJ_SUB];
#pragma unroll
    for (int k = 0; k < ADJ_SUB; ++k) {
      key[k] = adj_depth_key<DMAX>(P, c0 + k * ADJ_THREADS + threadIdx.x, lane_end);
      rank[k] = 0;
#pragma unroll
      for (int b = 0; b < ADJ_BINS; ++b) {
        const uint64_t m = __ballot(key[k] == b);
        if (key[k] == b) rank[k] = __popcll(m & lt);
        if ((threadIdx.x & 63) == 0) bins[(k * ADJ_WAVES + wave) * ADJ_BINS + b] = __popcll(m);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive scan, bins outermost
      int run = 0;
      for (int b = 0; b < ADJ_BINS; ++b)
        for (int r = 0; r < ADJ_SUB * ADJ_WAVES; ++r) {
          const int c = bins[r * ADJ_BINS + b];
          bins[r * ADJ_BINS + b] = run;
          run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ADJ_SUB; ++k)
      order[bins[(k * ADJ_WAVES + wave) * ADJ_BINS + key[k]] + rank[k]] = k * ADJ_THREADS + threadIdx.x;
    __syncthreads();
    for (int s = 0; s < ADJ_SUB; ++s) {
      const int64_t lane = c0 + order[s * ADJ_THREADS + threadIdx.x];
      if (lane >= lane_end) continue;
      const int64_t pix = P.spp > 1 ? lane / P.spp : lane;
      f3 g = f3{P.dLdI[3 * pix], P.dLdI[3 * pix + 1], P.dLdI[3 * pix + 2]};
      if (P.spp > 1) g = g / static_cast<float>(P.spp);
      if (g.x == 0.f && g.y == 0.f && g.z == 0.f) continue;
      const int info = P.cinfo[lane];
      if (info & 512) continue;  // the forward dropped this lane's non-finite radiance
      const Rng rng{0, lane_inc(static_cast<uint32_t>(lane))};
      int t0;
      float u0, v0;
      draw_cached(P, 0, lane, &t0, &u0, &v0);
      // camera ray origin: the camera position, or (orthographic) the film
      // point's spot on the camera plane, from the lane's first two draws
      f3 O0 = P.cam.pos, film = splat(0.f);
      if (P.cam.ortho) {
        Rng r0 = lane_rng(static_cast<uint32_t>(lane));
        const float a = r0.next(), b = r0.next();
        int pixel;
        film = film_point(P.cam, static_cast<uint32_t>(lane), P.spp, a, b, &pixel);
        const f3 fp = xform_pos(P.cam.r2c, film);
        O0 = xform_pos(P.cam.c2w, f3{fp.x, fp.y, 0.0f});
      }
      const Its its = reconstruct(sv, t0, u0, v0, O0, true);
      if (!its.valid) continue;
      // per-depth state recorded by the forward render: PCG state and throughput
      // at each vertex, and the vertices themselves (the extension hits)
      const int n = min(info & 255, DMAX);  // depths with light/BSDF work
      const bool light0 = (info & 256) != 0;  // camera vertex on a light, enable_light_visable
      // vertex k's cached hit: the camera hit (k = 0) or depth k-1's extension hit
      auto vtx = [&](int k, int* t, float* u, float* v) {
        if (k == 0) { *t = t0; *u = u0; *v = v0; return; }
        draw_cached(P, 1 + static_cast<int64_t>(k - 1) * P.Kmax + P.nL + P.nB, lane, t, u, v);
      };
      const f3 L = f3{P.cL[lane], P.cL[P.N + lane], P.cL[2 * P.N + lane]};
      const Its& cur = its;
      // enable_light_visable: L += T0 * emit at a light camera vertex; an SH9
      // light's radiance there depends on the camera origin through wi
      f3 gO0 = splat(0.f);
      if (light0) {
        const f3 gw0 = radiance_vjp<ENV>(P, hot, cur.light, cur, g);
        if (nonzero3(gw0)) gO0 = -vjp_normalize(cur.p - O0, -gw0);
      }
      // reverse sweep.  pend = adjoint of vertex d+1 gathered so far (its own
      // depth's terms and, through its wi, vertex d+2's); pendO = adjoint of
      // vertex d's point from vertex d+1's wi.
      f3 A = splat(0.f);  // adjoint of T_{d+1}
      VertexAdj pend = zero_adj();
      f3 pendO = splat(0.f);
      for (int d = n - 1; d >= 0; --d) {
        f3 O = O0;
        int ta, tb;
        float ua, va, ub, vb;
        if (d > 0) {
          vtx(d - 1, &ta, &ua, &va);
          O = reconstruct(sv, ta, ua, va, O0, true).p;
        }
        vtx(d, &tb, &ub, &vb);
        const Its vx = reconstruct(sv, tb, ub, vb, O, true);
        Rng r{P.crs[d * P.N + lane], rng.inc};
        const f3 Tsd =
            f3{P.cT[(3ll * d) * P.N + lane], P.cT[(3ll * d + 1) * P.N + lane], P.cT[(3ll * d + 2) * P.N + lane]};
        // A == 0: nothing after this depth reached L.  The throughput of an
        // extension whose sample has pdf 0 (cos = sqrt(u) = 0, or a grazing next
        // hit) is 0 / 0; the forward multiplies it only into terms it never
        // adds (else L is not finite and the lane is dropped), so its adjoint is
        // 0, not 0 * NaN.
        const bool tail = nonzero3(A);
        const DepthOut o = depth_step<ENV>(P, hot, vx, r, d, lane, true, g * Tsd, tail ? A * Tsd : splat(0.f));
        if (o.ext) {
          VertexAdj nx = pend;
          nx.p = nx.p + o.next.p;
          nx.n = nx.n + o.next.n;
          nx.area = nx.area + o.next.area;
          if (nonzero(nx)) {
            int tn;
            float un, vn;
            vtx(d + 1, &tn, &un, &vn);
            const Its nv = reconstruct(sv, tn, un, vn, vx.p, true);
            tri_vjp(P, tn, un, vn, flip_of(sv, nv), nx);
          }
        }
        pend = o.cur;
        pend.p = pend.p + pendO;
        pendO = o.gO;
        A = tail ? g * o.c + A * o.t : g * o.c;
      }
      // camera vertex: its own adjoints plus render_interior's L * J0 (J0 = A / detach(A))
      pend.area = pend.area + dot(g, L) / its.area;
      tri_vjp(P, t0, u0, v0, flip_of(sv, its), pend);
      // the camera ray origin, through the camera vertex's wi
      pendO = pendO + gO0;
      if (P.cam_acc && nonzero3(pendO)) (void)cam_origin_vjp(P.cam, pendO, film, P.cam.r2c, CAM_R, cacc);
    }
    __syncthreads();  // order[] and bins[] are rewritten by the next chunk
  }
  __syncthreads();
  adj_cache().flush(P.grad_v, P.grad_t);
  if (P.cam_acc) cam_acc_flush(cacc, P.cam_acc);
  for (int i = threadIdx.x; i < P.hot_n; i += blockDim.x) {
    const float v = hot[i];
    if (v == 0.f) continue;
    const int nl3 = 3 * P.sv.n_lights;
    if (i < nl3) {
      atomicAdd(P.grad_e + i, v);
    } else {
      const int o = (i - nl3) / 4, c = (i - nl3) % 4;
      const ObjectK& ob = P.sv.objs[o];
      if (c < 3) atomicAdd(P.grad_t + ob.alb_off + c, v);
      else atomicAdd(P.grad_t + ob.rough_off, v);
    }
  }
}

// one thread: camera matrix adjoints -> parameter gradients
__global__ void k_camera_grad(psdr_camera cam, const float* __restrict__ acc, float* __restrict__ grad) {
  if (blockIdx.x == 0 && threadIdx.x == 0) camera_param_vjp(cam, acc, grad);
}

inline int blocks(int64_t n, int t) { return static_cast<int>((n + t - 1) / t); }

}  // namespace

// The interior, primary-boundary and direct-boundary terms share only their
// (atomically accumulated) gradient outputs and read-only scene data, so they
// run side by side: the boundary terms on two auxiliary streams forked from and
// joined back into the caller's stream, each with its own scratch buffers.  The
// parts of the boundary terms that do not depend on dL/dI are normally issued by
// the forward render (boundary_prepare_async) and overlap it; the backward then
// only forms the dL/dI-weighted scatters.

// Joins open auxiliary-stream work into `st`.  Under stream capture only work
// issued in the same capture can be joined (a captured stream may not wait on an
// event recorded outside it); work left open by an eager renderD whose backward
// never ran must be finished before a capture begins.
int aux_join(psdr_scene* s, hipStream_t st) {
  if (!s->ws->aux.fork || !s->ws->aux_open) return PSDR_OK;
  if (capture_active() && !s->ws->aux_captured) {
    set_error(PSDR_ERR_CAPTURE, "boundary work of an eager renderD without backward is still open: run its "
                                "backward before capturing");
    return PSDR_ERR_CAPTURE;
  }
  for (int i = 0; i < 2; ++i) {
    PSDR_HIP_CHECK(hipEventRecord(s->ws->aux.join[i], s->ws->aux.st[i]));
    PSDR_HIP_CHECK(hipStreamWaitEvent(st, s->ws->aux.join[i], 0));
  }
  s->ws->aux_open = false;
  return PSDR_OK;
}

static int aux_fork(psdr_scene* s, hipStream_t st) {
  int rc = s->ws->aux.ensure();
  if (rc) return rc;
  if (s->ws->aux_open && capture_active() != s->ws->aux_captured) {
    // open work from outside this capture (or a capture left unjoined)
    rc = aux_join(s, st);
    if (rc) return rc;
  }
  PSDR_HIP_CHECK(hipEventRecord(s->ws->aux.fork, st));
  for (int i = 0; i < 2; ++i) PSDR_HIP_CHECK(hipStreamWaitEvent(s->ws->aux.st[i], s->ws->aux.fork, 0));
  s->ws->aux_open = true;
  s->ws->aux_captured = capture_active();
  return PSDR_OK;
}

int boundary_prepare_async(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam) {
  s->ws->cur->bprep = false;
  if (s->n_edges == 0 || s->plain || !s->ws->aux.fork) return PSDR_OK;
  for (int i = 0; i < 2; ++i) PSDR_HIP_CHECK(hipStreamWaitEvent(s->ws->aux.st[i], s->ws->aux.fork, 0));
  int rc = PSDR_OK;
  // the direct chain first: a few launches, one of them long, so it is queued
  // on the GPU at once; the primary chain is a long run of launches after it
  if (cfg.spp_secondary_edge > 0) rc = boundary_direct_prepare(s, cfg, cam, s->ws->aux.st[1]);
  if (!rc && cfg.spp_primary_edge > 0) rc = boundary_primary_prepare(s, cfg, cam, s->ws->aux.st[0]);
  s->ws->aux_open = true;
  s->ws->aux_captured = capture_active();
  if (rc) return rc;
  s->ws->cur->bprep = true;
  return PSDR_OK;
}

static int camera_grad_begin(psdr_scene* s, float* grad_cam, hipStream_t st, float** acc) {
  *acc = nullptr;
  if (!grad_cam) return PSDR_OK;
  if (s->ws->cam_acc.ensure(4ull * CAM_ACC)) return grow_failed();
  *acc = s->ws->cam_acc.as<float>();
  return fill_f32(*acc, CAM_ACC, 0.f, st);
}

static int camera_grad_end(const psdr_camera& camd, const float* acc, float* grad_cam, hipStream_t st) {
  if (!grad_cam) return PSDR_OK;
  hipLaunchKernelGGL(k_camera_grad, dim3(1), dim3(1), 0, st, camd, acc, grad_cam);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

int render_backward(psdr_scene* s, const psdr_config& cfg, const psdr_camera& camd, const float* dL_dI,
                    float* grad_v, float* grad_t, float* grad_e, float* grad_cam, int terms, hipStream_t st) {
  const bool ti = (terms & PSDR_TERM_INTERIOR) != 0;
  const bool tp = (terms & PSDR_TERM_PRIMARY) != 0 && s->n_edges > 0 && cfg.spp_primary_edge > 0;
  const bool td = (terms & PSDR_TERM_DIRECT) != 0 && s->n_edges > 0 && cfg.spp_secondary_edge > 0;
  if (s->ws->stats.ensure(4ull * (3 * static_cast<size_t>(s->n_vertices) + s->n_texels + 3 * s->n_lights + 64)))
    return grow_failed();
  float* gv = grad_v ? grad_v : s->ws->stats.as<float>();
  RenderSlot* slot = s->ws->find_slot(s, cfg, camd);
  if (!slot) {
    set_error(PSDR_ERR_STATE, "render_backward: no matching forward render with keep=1");
    return PSDR_ERR_STATE;
  }
  s->ws->cur = slot;
  float* acc;
  int rc0 = camera_grad_begin(s, grad_cam, st, &acc);
  if (rc0) return rc0;
  if (s->plain) {
    int rc = ti ? interior_backward(s, cfg, camd, dL_dI, grad_v, grad_t, grad_e, acc, st) : PSDR_OK;
    if (!rc && tp) rc = boundary_primary_prepare(s, cfg, camd, st);
    if (!rc && tp) rc = boundary_primary_finish(s, cfg, camd, dL_dI, gv, acc, st);
    if (!rc && td) rc = boundary_direct_prepare(s, cfg, camd, st);
    if (!rc && td) rc = boundary_direct_finish(s, cfg, camd, dL_dI, gv, st);
    if (!rc) rc = camera_grad_end(camd, acc, grad_cam, st);
    return rc;
  }
  const bool prepared = slot->bprep;
  int rc = (tp || td) ? aux_fork(s, st) : PSDR_OK;
  if (rc) return rc;
  // launch order by host enqueue cost: the single interior launch first
  int rci = ti ? interior_backward(s, cfg, camd, dL_dI, grad_v, grad_t, grad_e, acc, st) : PSDR_OK;
  int rcd = PSDR_OK, rcp = PSDR_OK;
  if (td) {
    if (!prepared) rcd = boundary_direct_prepare(s, cfg, camd, s->ws->aux.st[1]);
    if (!rcd) rcd = boundary_direct_finish(s, cfg, camd, dL_dI, gv, s->ws->aux.st[1]);
  }
  if (tp) {
    if (!prepared) rcp = boundary_primary_prepare(s, cfg, camd, s->ws->aux.st[0]);
    if (!rcp) rcp = boundary_primary_finish(s, cfg, camd, dL_dI, gv, acc, s->ws->aux.st[0]);
  }
  // join even after an error so no work is left running unordered (also a
  // prepass of this forward whose terms this call does not form)
  const int rcj = aux_join(s, st);
  if (rci) return rci;
  if (rcp) return rcp;
  if (rcd) return rcd;
  if (rcj) return rcj;
  return camera_grad_end(camd, acc, grad_cam, st);
}

int interior_backward(psdr_scene* s, const psdr_config& cfg, const psdr_camera& camd, const float* dL_dI,
                      float* grad_v, float* grad_t, float* grad_e, float* cam_acc, hipStream_t st) {
  AdjParams P;
  P.sv = s->view();
  P.cam = make_camera(camd);
  P.D = static_cast<int>(cfg.max_bounce);
  P.nL = cfg.mis_light_samples;
  P.nB = cfg.mis_bsdf_samples;
  P.elv = cfg.enable_light_visable ? 1 : 0;
  P.spp = cfg.spp_interior;
  const RenderSlot& R = *s->ws->cur;
  P.Kmax = R.K;
  P.N = static_cast<int64_t>(R.N);
  const int64_t slots = 1 + static_cast<int64_t>(R.D) * R.K;
  P.ctri = R.hit_cache.as<int>();
  P.cu = reinterpret_cast<const float*>(P.ctri + slots * P.N);
  P.cv = P.cu + slots * P.N;
  const int64_t Dr = std::max(R.D, 1);
  P.crs = R.replay.as<uint64_t>();
  P.cT = reinterpret_cast<const float*>(P.crs + Dr * P.N);
  P.cL = P.cT + 3 * Dr * P.N;
  P.cinfo = reinterpret_cast<const int*>(P.cL + 3 * P.N);
  P.dLdI = dL_dI;
  // scratch for absent gradient buffers
  if (s->ws->stats.ensure(4ull * (3 * static_cast<size_t>(s->n_vertices) + s->n_texels + 3 * s->n_lights + 64)))
    return grow_failed();
  float* sink = s->ws->stats.as<float>();
  P.grad_v = grad_v ? grad_v : sink;
  P.grad_t = grad_t ? grad_t : sink;
  P.grad_e = grad_e ? grad_e : sink;
  P.cam_acc = cam_acc;
  const int hot = 3 * s->n_lights + 4 * s->n_objects;
  P.hot_n = hot <= HOT_MAX ? hot : 0;
  const int grid = std::max(1, std::min(blocks(P.N, ADJ_THREADS), num_cus() * tuning().adjoint_bpc.load()));
  const bool env = s->n_bvh < s->n_tris;  // an environment light's cube
  if (P.D <= 4) {
    if (env) hipLaunchKernelGGL((k_adjoint<4, true>), dim3(grid), dim3(ADJ_THREADS), 0, st, P);
    else hipLaunchKernelGGL((k_adjoint<4, false>), dim3(grid), dim3(ADJ_THREADS), 0, st, P);
  } else if (P.D <= 8) {
    if (env) hipLaunchKernelGGL((k_adjoint<8, true>), dim3(grid), dim3(ADJ_THREADS), 0, st, P);
    else hipLaunchKernelGGL((k_adjoint<8, false>), dim3(grid), dim3(ADJ_THREADS), 0, st, P);
  } else {
    if (env) hipLaunchKernelGGL((k_adjoint<64, true>), dim3(grid), dim3(ADJ_THREADS), 0, st, P);
    else hipLaunchKernelGGL((k_adjoint<64, false>), dim3(grid), dim3(ADJ_THREADS), 0, st, P);
  }
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

}  // namespace psdr
