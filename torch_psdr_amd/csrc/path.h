// One depth of trace2light (src/path_integrator.cc:80-186), split into the
// parts the wavefront kernel (render.hip k_shade) and the fused primary-edge
// preprocessing kernel (boundary.hip k_prim_pre) share: the ray each sample
// of a depth makes, and what its hit adds to the path.  Both callers draw the
// same random numbers in the same order and pass the hit of the same ray, so
// they compute the same values.
#pragma once

#include "shading.h"

namespace psdr {

// Next-event estimation, ray (src/path_integrator.cc:103-110): a light picked
// by power, a point on it, the direction to it and its distance; a NaN
// direction when the picked object emits nothing.
__device__ __forceinline__ void light_sample_ray(const SceneView& sv, f3 p, float ru, float rv, f3* dir, float* tl) {
  float pdf_choose;
  const int obj = discrete_sample(sv.obj_pmf, sv.obj_cmf, sv.n_objects, &ru, &pdf_choose);
  const int l = sv.objs[obj].light;
  *dir = splat(__builtin_nanf(""));
  *tl = F_INF;
  if (l >= 0) {
    const LightPoint lp = sample_light_point(sv, l, ru, rv);
    *dir = normalize(lp.p - p);
    *tl = norm(lp.p - p);
  }
}

// Next-event estimation, contribution at vertex ip (src/path_integrator.cc:103-131)
// of the light-sample ray with hit (ht, hu, hv): L += T * w * value.
__device__ __forceinline__ void light_sample_add(const SceneView& sv, const Its& ip, float ru, float rv, int ht,
                                                 float hu, float hv, uint32_t nL, uint32_t nB, int ad_jac, f3 T,
                                                 f3* L) {
  float pdf_choose;
  const int obj = discrete_sample(sv.obj_pmf, sv.obj_cmf, sv.n_objects, &ru, &pdf_choose);
  const int l = sv.objs[obj].light;
  if (l < 0) return;
  const LightPoint lp = sample_light_point(sv, l, ru, rv);
  const float pdf_light = lp.pdf * pdf_choose;
  const f3 dir = normalize(lp.p - ip.p);
  const Its itl = reconstruct(sv, ht, hu, hv, ip.p, true);
  const bool is_hit = itl.valid && itl.is_light && sqnorm(itl.p - lp.p) < REAL_EPS;
  if (!is_hit) return;
  const f3 emit = light_radiance(sv, l, itl);
  const f3 bsdf = mat_eval(sv, ip, dir);
  const float tmp = dot(itl.n, -dir) / itl.d2;
  const float G = dot(ip.n, dir) * tmp;
  f3 value = (emit * G) * bsdf;
  if (ad_jac) value = value * (itl.area / itl.area);
  value = value / pdf_light;
  float weight = 1.0f / static_cast<float>(nL);
  if (nB > 0) weight = weight * mis_weight(pdf_light, mat_pdf(ip, dir) * tmp, nL, nB);
  *L = *L + (T * weight) * value;
}

// BSDF sampling with MIS, contribution at vertex ip (src/path_integrator.cc:134-160)
// of the BSDF-sampled ray with hit (ht, hu, hv): L += T * w * value when the
// hit is an emitting triangle.
__device__ __forceinline__ void bsdf_sample_add(const SceneView& sv, const Its& ip, float ru, float rv, int ht,
                                                float hu, float hv, uint32_t nL, uint32_t nB, int ad_jac, f3 T,
                                                f3* L) {
  const Sampled smp = mat_sample(sv, ip, ru, rv);
  const Its itl = reconstruct(sv, ht, hu, hv, ip.p, true);
  if (!(itl.valid && itl.is_light)) return;
  const int l = itl.light;
  const f3 emit = light_radiance(sv, l, itl);
  const float tmp = dot(itl.n, -smp.wo) / itl.d2;
  const float G = dot(ip.n, smp.wo) * tmp;
  const float pdf_bsdf = smp.pdf * tmp;
  f3 value = (emit * G) * smp.bsdf;
  if (ad_jac) value = value * (itl.area / itl.area);
  value = value / pdf_bsdf;
  float weight = 1.0f / static_cast<float>(nB);
  if (nL > 0) {
    const float pdf_light = (sv.light_power[l] / *sv.lights_total) * (1.0f / sv.light_area[l]);
    weight = weight * mis_weight(pdf_bsdf, pdf_light, nB, nL);
  }
  *L = *L + (T * weight) * value;
}

// Extension (src/path_integrator.cc:163-181): the next vertex from the
// extension ray's hit (ht, hu, hv) at vertex ip; T takes the sample's weight
// and O the vertex it leaves from.  Returns whether the path goes on.
__device__ __forceinline__ bool extend_path(const SceneView& sv, const Its& ip, float ru, float rv, int ht, float hu,
                                            float hv, int ad_jac, Its* its, f3* T, f3* O) {
  const Sampled smp = mat_sample(sv, ip, ru, rv);
  *its = reconstruct(sv, ht, hu, hv, ip.p, true);
  const bool alive = smp.valid && its->valid;
  if (alive) {
    const float tmp = dot(its->n, -smp.wo) / its->d2;
    const float G = dot(ip.n, smp.wo) * tmp;
    const float pdf = smp.pdf * tmp;
    f3 f = smp.bsdf;
    if (ad_jac) f = f * (its->area / its->area);
    *T = *T * ((f * G) / pdf);
    *O = ip.p;
  }
  return alive;
}

}  // namespace psdr
