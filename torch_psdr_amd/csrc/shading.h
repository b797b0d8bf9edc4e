// Device-side scene access: intersection reconstruction, textures, materials
// and area-light sampling.  Each routine restates the reference expression it
// cites, in the same fp32 operation order.
#pragma once

#include "scene.h"

namespace psdr {

struct Its {
  bool valid, is_light;
  f3 p, n, t, wi;
  float d2;
  f2 uv;
  float area;
  int tri, obj, light;
};

__device__ __forceinline__ f3 xyz4(float4 a) { return f3{a.x, a.y, a.z}; }

// The scene's small tables (objects, lights, the light-pick and light-triangle
// samplers, the emitting triangles' leaves) staged in a workgroup's LDS.  A
// shading step looks them up several times in a chain of dependent reads (the
// hit's object, then the picked light, its triangle sampler, the sampled
// triangle); from LDS each read costs a fraction of a cached global load's
// latency.  Values are copied bit for bit, so results do not change.
constexpr int LDS_OBJS = 16, LDS_LIGHTS = 8, LDS_LIGHT_TRIS = 64;
struct LdsTables {
  ObjectK objs[LDS_OBJS];
  LightK lights[LDS_LIGHTS];
  float obj_pmf[LDS_OBJS], obj_cmf[LDS_OBJS];
  float light_emit[3 * LDS_LIGHTS], light_power[LDS_LIGHTS], light_area[LDS_LIGHTS], lights_total[4];
  float ltri_pmf[LDS_LIGHT_TRIS], ltri_cmf[LDS_LIGHT_TRIS];
  float4 seek[3 * SEEK_MAX];
};
// whether a scene's tables fit LdsTables (host side: picks the kernel instance)
inline bool lds_tables_fit(const SceneView& v) {
  return v.n_objects <= LDS_OBJS && v.n_lights <= LDS_LIGHTS && v.n_light_tris <= LDS_LIGHT_TRIS &&
         v.n_seek <= SEEK_MAX;
}
template <class T>
__device__ __forceinline__ void lds_copy_words(T* dst, const T* src, int n) {
  static_assert(sizeof(T) % 4 == 0, "word copy");
  int* d = reinterpret_cast<int*>(dst);
  const int* s = reinterpret_cast<const int*>(src);
  const int words = n * static_cast<int>(sizeof(T) / 4);
  for (int i = threadIdx.x; i < words; i += blockDim.x) d[i] = s[i];
}
// Every thread of the block calls this (it ends with a barrier); v then reads
// the staged tables.  The caller checked lds_tables_fit.
__device__ __forceinline__ void stage_tables(LdsTables& t, SceneView* v) {
  lds_copy_words(t.objs, v->objs, v->n_objects);
  lds_copy_words(t.lights, v->lights, v->n_lights);
  lds_copy_words(t.obj_pmf, v->obj_pmf, v->n_objects);
  lds_copy_words(t.obj_cmf, v->obj_cmf, v->n_objects);
  lds_copy_words(t.light_emit, v->light_emit, 3 * v->n_lights);
  lds_copy_words(t.light_power, v->light_power, v->n_lights);
  lds_copy_words(t.light_area, v->light_area, v->n_lights);
  lds_copy_words(t.lights_total, v->lights_total, 1);
  lds_copy_words(t.ltri_pmf, v->light_tri_pmf, v->n_light_tris);
  lds_copy_words(t.ltri_cmf, v->light_tri_cmf, v->n_light_tris);
  if (v->n_seek > 0) lds_copy_words(t.seek, v->seek_leaf, 3 * v->n_seek);
  __syncthreads();
  v->objs = t.objs;
  v->lights = t.lights;
  v->obj_pmf = t.obj_pmf;
  v->obj_cmf = t.obj_cmf;
  v->light_emit = t.light_emit;
  v->light_power = t.light_power;
  v->light_area = t.light_area;
  v->lights_total = t.lights_total;
  v->light_tri_pmf = t.ltri_pmf;
  v->light_tri_cmf = t.ltri_cmf;
  v->seek_leaf = t.seek;
}

// Scene::hit reconstruction (src/scene.cc:106-161) from (tri, u, v) and the ray
// origin O.
__device__ __forceinline__ Its reconstruct(const SceneView& sv, int tri, float u, float v, f3 O, bool valid_in) {
  Its it;
  it.valid = valid_in && tri >= 0;
  it.is_light = false;
  it.tri = tri;
  it.obj = -1;
  it.light = -1;
  if (!it.valid) {
    it.p = it.n = it.t = it.wi = splat(0.f);
    it.d2 = 0.f;
    it.uv = f2{0.f, 0.f};
    it.area = 0.f;
    return it;
  }
  const float4 g0 = sv.tri_geom[5 * tri], g1 = sv.tri_geom[5 * tri + 1], g2 = sv.tri_geom[5 * tri + 2];
  const float4 g3 = sv.tri_geom[5 * tri + 3], g4 = sv.tri_geom[5 * tri + 4];
  const f3 P0 = xyz4(g0), P1 = xyz4(g1), P2 = xyz4(g2), N = xyz4(g3), T = xyz4(g4);
  it.area = g0.w;
  it.p = (P0 + u * (P1 - P0)) + v * (P2 - P0);
  const float* uv = sv.tri_uv + 6 * tri;
  it.uv.x = (uv[0] + u * (uv[2] - uv[0])) + v * (uv[4] - uv[0]);
  it.uv.y = (uv[1] + u * (uv[3] - uv[1])) + v * (uv[5] - uv[1]);
  const f3 wu = it.p - O;
  it.wi = -normalize(wu);
  it.d2 = sqnorm(wu);
  const bool front = dot(it.wi, N) > 0.f;
  it.n = front ? N : -N;
  it.t = front ? T : -T;
  it.obj = sv.tri_obj[tri];
  it.light = sv.objs[it.obj].light;
  it.is_light = it.light >= 0;
  return it;
}

// UVTexture::at_nearest (src/texture.h:65-74); the Python wrapper flips rows,
// so data row v0 is image row H-1-v0 of the caller's (H, W, C) tensor.
__device__ __forceinline__ int texel_index(int h, int w, f2 uv) {
  const float u = uv.x * static_cast<float>(w) - 0.5f;
  const float v = uv.y * static_cast<float>(h) - 0.5f;
  const float u0 = fminf(fmaxf(rintf(u), 0.f), static_cast<float>(w - 1));
  const float v0 = fminf(fmaxf(rintf(v), 0.f), static_cast<float>(h - 1));
  return (h - 1 - static_cast<int>(v0)) * w + static_cast<int>(u0);
}
// float offset of the 3-channel texel for uv
__device__ __forceinline__ int tex3_offset(int off, int h, int w, f2 uv) {
  return h == 0 ? off : off + 3 * texel_index(h, w, uv);
}
__device__ __forceinline__ int tex1_offset(int off, int h, int w, f2 uv) {
  return h == 0 ? off : off + texel_index(h, w, uv);
}
__device__ __forceinline__ f3 tex3(const float* tex, int o) { return f3{tex[o], tex[o + 1], tex[o + 2]}; }

// Second-order real spherical harmonics Y_0..Y_8 of a unit direction
// (SH9Light::eval_impl_, src/emitter.cc:251-268): the constants are the fp32
// square roots of 1/(4 pi), 3/(4 pi), 15/(4 pi), 5/(16 pi), 15/(16 pi).
struct Sh9Const {
  float c0, c1, c2, c3, c4;
};
__device__ __forceinline__ Sh9Const sh9_const() {
  return Sh9Const{sqrtf(REAL_INV_4PI), sqrtf(0.75f * REAL_INV_PI), sqrtf(3.75f * REAL_INV_PI),
                  sqrtf(0.3125f * REAL_INV_PI), sqrtf(0.9375f * REAL_INV_PI)};
}
struct Sh9Basis {
  float y[9];
};
__device__ __forceinline__ Sh9Basis sh9_basis(f3 w) {
  const Sh9Const c = sh9_const();
  const float xx = w.x * w.x, yy = w.y * w.y, zz = w.z * w.z;
  const float xy = w.x * w.y, yz = w.y * w.z, zx = w.z * w.x;
  Sh9Basis b;
  b.y[0] = c.c0;
  b.y[1] = c.c1 * w.y;
  b.y[2] = c.c1 * w.z;
  b.y[3] = c.c1 * w.x;
  b.y[4] = c.c2 * xy;
  b.y[5] = c.c2 * yz;
  b.y[6] = c.c3 * ((2.0f * zz - xx) - yy);
  b.y[7] = c.c2 * zx;
  b.y[8] = c.c4 * (xx - yy);
  return b;
}

// Radiance of light l at its hit `it` (Light::eval): an AreaLight's constant
// emission, two-sided (src/emitter.cc:16-23); an EnvLight's cube map at the
// hit's uv (DiffuseLightMaterial::emit, src/diffuse_material.h:53-58,
// src/emitter.cc:133-140); an SH9Light's sum_k C_k Y_k(wi) over the 9 RGB
// coefficient rows (src/emitter.cc:269-279), summed in k order.
__device__ __forceinline__ f3 light_radiance(const SceneView& sv, int l, const Its& it) {
  const ObjectK& ob = sv.objs[sv.lights[l].obj];
  if (ob.mat == PSDR_MAT_ENV_MAP) return tex3(sv.texels, tex3_offset(ob.alb_off, ob.alb_h, ob.alb_w, it.uv));
  if (ob.mat == PSDR_MAT_ENV_SH9) {
    const Sh9Basis b = sh9_basis(it.wi);
    const float* c = sv.texels + ob.alb_off;
    f3 r = f3{c[0], c[1], c[2]} * b.y[0];
#pragma unroll
    for (int k = 1; k < 9; ++k) r = r + f3{c[3 * k], c[3 * k + 1], c[3 * k + 2]} * b.y[k];
    return r;
  }
  return f3{sv.light_emit[3 * l], sv.light_emit[3 * l + 1], sv.light_emit[3 * l + 2]};
}

// DiffuseBsdf shading normal (src/diffuse_bsdf.cc:13-19, 61-68): the
// geometric normal, or with a normal map normalize(TBN (2 n - 1)) where TBN has
// the columns T, B = Ng x T, Ng (the matrix-vector product summed in that order).
__device__ __forceinline__ f3 normal_map_value(const SceneView& sv, const ObjectK& ob, const Its& it) {
  const int o = tex3_offset(ob.nrm_off, ob.nrm_h, ob.nrm_w, it.uv);
  const f3 tn = tex3(sv.texels, o);
  const f3 v = tn * 2.0f - splat(1.0f);
  const f3 B = cross(it.n, it.t);
  return normalize((it.t * v.x + B * v.y) + it.n * v.z);
}
__device__ __forceinline__ f3 shading_normal(const SceneView& sv, const ObjectK& ob, const Its& it) {
  return (ob.mat == PSDR_MAT_DIFFUSE_BSDF && ob.nrm_off >= 0) ? normal_map_value(sv, ob, it) : it.n;
}

// DiffuseBsdfMaterial value (src/diffuse_bsdf.cc:42-75) for shading normal N
__device__ __forceinline__ f3 diffuse_bsdf_value(f3 color, float sigma, f3 N, const Its& it, f3 L) {
  const float div = 1.0f / (REAL_PI + ((3.0f * REAL_PI - 4.0f) / 6.0f) * sigma);
  const float a = div;
  const float b = sigma * div;
  float NdotL = dot(N, L);
  if (NdotL < 0.f) NdotL = 0.f;
  float NdotV = dot(N, it.wi);
  if (NdotV < 0.f) NdotV = 0.f;
  const float VdotL = dot(it.wi, L);
  float t = VdotL - NdotL * NdotV;
  if (t > 0.f) t = (NdotL >= NdotV ? NdotL : NdotV) + REAL_EPS;
  return color * (a + b * t);
}

// Material::eval (diffuse_material.cc:29-34, diffuse_bsdf.cc:42-75)
__device__ __forceinline__ f3 mat_eval(const SceneView& sv, const Its& it, f3 wo) {
  const ObjectK& ob = sv.objs[it.obj];
  const f3 color = tex3(sv.texels, tex3_offset(ob.alb_off, ob.alb_h, ob.alb_w, it.uv));
  if (ob.mat == PSDR_MAT_DIFFUSE_BSDF) {
    const float sigma = sv.texels[tex1_offset(ob.rough_off, ob.rough_h, ob.rough_w, it.uv)];
    return diffuse_bsdf_value(color, sigma, shading_normal(sv, ob, it), it, wo);
  }
  return color * REAL_INV_PI;
}

// Material::pdf
__device__ __forceinline__ float mat_pdf(const Its& it, f3 wo) { return dot(it.n, wo) * REAL_INV_PI; }

struct Sampled {
  f3 wo, bsdf;
  float pdf, cos_t, u, v;
  bool valid;
};
// Material::sample (diffuse_material.cc:3-27, diffuse_bsdf.cc:3-40)
__device__ __forceinline__ Sampled mat_sample(const SceneView& sv, const Its& it, float ru, float rv) {
  Sampled s;
  const ObjectK& ob = sv.objs[it.obj];
  const f3 B = cross(it.n, it.t);
  const f3 N = shading_normal(sv, ob, it);
  const float cos_t = safe_sqrt(ru);
  const float sin_t = safe_sqrt(1.0f - ru);
  const float phi = rv * REAL_2PI;
  float sp, cp;
  psdr_sincos(phi, &sp, &cp);
  const float u = cp * sin_t;
  const float v = sp * sin_t;
  s.wo = normalize((u * it.t + v * B) + cos_t * N);
  s.pdf = cos_t * REAL_INV_PI;
  s.cos_t = cos_t;
  s.u = u;
  s.v = v;
  s.valid = it.valid;
  const f3 color = tex3(sv.texels, tex3_offset(ob.alb_off, ob.alb_h, ob.alb_w, it.uv));
  if (ob.mat == PSDR_MAT_DIFFUSE_BSDF) {
    s.valid = s.valid && dot(it.n, it.wi) > 0.f;
    const float sigma = sv.texels[tex1_offset(ob.rough_off, ob.rough_h, ob.rough_w, it.uv)];
    s.bsdf = diffuse_bsdf_value(color, sigma, N, it, s.wo);
  } else {
    s.bsdf = color * REAL_INV_PI;
  }
  return s;
}

// AreaLight::sample_light_point_ (src/emitter.cc:30-59).  Returns the light
// triangle (global id) and the barycentric weights used for the point.
struct LightPoint {
  f3 p;
  float pdf;
  int tri;
  float bu, bv;
};
__device__ __forceinline__ LightPoint sample_light_point(const SceneView& sv, int l, float ru, float rv) {
  const LightK L = sv.lights[l];
  float pt;
  const int k = discrete_sample(sv.light_tri_pmf + L.pmf_off, sv.light_tri_cmf + L.pmf_off, L.tri_count, &ru, &pt);
  const int tri = L.tri_begin + k;
  const float4 g0 = sv.tri_geom[5 * tri], g1 = sv.tri_geom[5 * tri + 1], g2 = sv.tri_geom[5 * tri + 2];
  const f3 P0 = xyz4(g0), P1 = xyz4(g1), P2 = xyz4(g2);
  const float svv = safe_sqrt(rv);
  const float u = ru * svv;
  const float v = 1.0f - svv;
  LightPoint lp;
  lp.pdf = 1.0f / sv.light_area[l];
  lp.p = (P0 + (P1 - P0) * u) + (P2 - P0) * v;
  lp.tri = tri;
  lp.bu = u;
  lp.bv = v;
  return lp;
}

// PathIntegrator::mis_weight (src/path_integrator.cc:3-9)
__device__ __forceinline__ float mis_weight(float w1, float w2, uint32_t n1, uint32_t n2) {
  const float a = static_cast<float>(n1) * w1;
  const float b = static_cast<float>(n2) * w2;
  const float sa = a * a, sb = b * b;
  return sa / (sa + sb);
}

// Workgroup-aggregated slot allocation: one atomic per workgroup (ballot
// counts per wave, their prefix in LDS).  Atomic adds to one address are
// serialised by the memory system (about 11 ns each on MI355X, measured), so
// a queue filled with one atomic per wave by tens of thousands of waves spends
// most of its time there.  Every thread of the block must call it (it holds
// three barriers); lds: (blockDim / 64 + 1) ints.
// Slots are dense: wave w's lanes follow those of waves 0 .. w-1.
__device__ __forceinline__ int block_alloc(bool want, int* counter, int* lds) {
  const uint64_t m = __ballot(want);
  const int lane = __lane_id(), w = static_cast<int>(threadIdx.x) >> 6, nw = static_cast<int>(blockDim.x) >> 6;
  if (lane == 0) lds[w] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int k = 0; k < nw; ++k) {
      const int c = lds[k];
      lds[k] = s;
      s += c;
    }
    lds[nw] = s != 0 ? atomicAdd(counter, s) : 0;
  }
  __syncthreads();
  const uint64_t lt = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
  const int slot = lds[nw] + lds[w] + __popcll(m & lt);
  __syncthreads();  // lds is reused by the next call
  return slot;
}

}  // namespace psdr
