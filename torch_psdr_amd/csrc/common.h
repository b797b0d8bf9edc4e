// Shared device/host math for the gfx950 path tracer.
//
// All arithmetic is plain IEEE fp32 (+ - * / sqrt) and the library is compiled
// with -ffp-contract=off, so every value below is reproducible bit-for-bit by a
// CPU restatement evaluating the same expressions in the same order.  The
// expression orders follow the reference's C++ (cited per function).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PSDR_HD __host__ __device__ __forceinline__

namespace psdr {

constexpr float REAL_EPS = 1e-5f;                    // src/types.h:114
constexpr float RAY_TMIN = 1e-3f;                    // src/cuda/device_programs.cu:8
constexpr float REAL_PI = 3.14159274101257324f;      // src/types.h:118
constexpr float REAL_2PI = 6.28318548202514648f;     // src/types.h:119
constexpr float REAL_4PI = 12.5663709640502930f;     // src/types.h:120
constexpr float REAL_INV_4PI = 0.0795774683356285f;  // src/types.h:123
constexpr float REAL_INV_PI = 0.318309873342514038f; // src/types.h:121
constexpr float F_INF = __builtin_huge_valf();

struct f3 {
  float x, y, z;
};
struct f2 {
  float x, y;
};

PSDR_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
PSDR_HD f3 splat(float a) { return f3{a, a, a}; }
PSDR_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PSDR_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PSDR_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
PSDR_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PSDR_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PSDR_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
PSDR_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
PSDR_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
PSDR_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
PSDR_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PSDR_HD float sqnorm(f3 a) { return dot(a, a); }
PSDR_HD float norm(f3 a) { return sqrtf(sqnorm(a)); }
// enoki::normalize restated as v * (1 / sqrt(|v|^2))
PSDR_HD f3 normalize(f3 a) {
  const float inv = 1.0f / sqrtf(sqnorm(a));
  return a * inv;
}
PSDR_HD float safe_sqrt(float x) { return sqrtf(fmaxf(x, 0.f)); }
PSDR_HD float get(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// ---------------------------------------------------------------------------
// sin / cos: Cody-Waite reduction by pi/2 (split 1.5703125 + 4.8375e-4 +
// 7.55e-8) and Taylor polynomials (coefficients 1/n!) on [-pi/4, pi/4].
// ---------------------------------------------------------------------------
PSDR_HD void psdr_sincos(float x, float* s, float* c) {
  const float k = rintf(x * 0.636619746685028076f);
  float r = x - k * 1.5703125f;
  r = r - k * 4.83751296997070312e-4f;
  r = r - k * 7.54978995489188216e-8f;
  const int q = static_cast<int>(k) & 3;
  const float r2 = r * r;
  float ps = -2.50521079e-8f;
  ps = ps * r2 + 2.75573188e-6f;
  ps = ps * r2 - 1.98412701e-4f;
  ps = ps * r2 + 8.33333377e-3f;
  ps = ps * r2 - 1.66666672e-1f;
  const float sp = r + (r * r2) * ps;
  float pc = 2.08767559e-9f;
  pc = pc * r2 - 2.75573200e-7f;
  pc = pc * r2 + 2.48015876e-5f;
  pc = pc * r2 - 1.38888892e-3f;
  pc = pc * r2 + 4.16666679e-2f;
  const float cp = (1.0f - 0.5f * r2) + (r2 * r2) * pc;
  if (q == 0) {
    *s = sp;
    *c = cp;
  } else if (q == 1) {
    *s = cp;
    *c = -sp;
  } else if (q == 2) {
    *s = -sp;
    *c = -cp;
  } else {
    *s = -cp;
    *c = sp;
  }
}

// ---------------------------------------------------------------------------
// PCG32 XSH-RR (O'Neill 2014) as used by Enoki's PCG32, TEA-seeded per lane
// (src/sampler.h:11-17, 70-80).
// ---------------------------------------------------------------------------
constexpr uint64_t PCG32_DEFAULT_STATE = 0x853c49e6748fea9bULL;
constexpr uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;

PSDR_HD uint32_t pcg_output(uint64_t old) {
  const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = static_cast<uint32_t>(old >> 59u);
  return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
}
PSDR_HD float u32_to_unit(uint32_t u) {
  const uint32_t b = (u >> 9) | 0x3f800000u;
  return __builtin_bit_cast(float, b) - 1.0f;
}
struct Rng {
  uint64_t state, inc;
  PSDR_HD uint32_t next_u32() {
    const uint64_t old = state;
    state = old * PCG32_MULT + inc;
    return pcg_output(old);
  }
  PSDR_HD float next() { return u32_to_unit(next_u32()); }
};
PSDR_HD uint64_t tea64(uint32_t v0, uint32_t v1) {
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    sum += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return static_cast<uint64_t>(v0) + (static_cast<uint64_t>(v1) << 32);
}
PSDR_HD uint64_t lane_inc(uint32_t lane) {
  const uint32_t s = static_cast<uint32_t>(static_cast<uint64_t>(lane) + PCG32_DEFAULT_STATE);
  return (tea64(lane, s) << 1) | 1u;
}
PSDR_HD Rng lane_rng(uint32_t lane) {
  const uint32_t s = static_cast<uint32_t>(static_cast<uint64_t>(lane) + PCG32_DEFAULT_STATE);
  const uint64_t initstate = tea64(s, lane);
  Rng r;
  r.state = 0;
  r.inc = (tea64(lane, s) << 1) | 1u;
  r.next_u32();
  r.state += initstate;
  r.next_u32();
  return r;
}

// ---------------------------------------------------------------------------
// Ray / triangle: Moller-Trumbore, fixed operation order; u, v are the weights of
// p1, p2 (device_programs.cu:45).  Hit iff t >= tmin (1e-3).  Split in two so
// that traversal can reject most triangles from the terms alone, without the
// division: tri_terms gives det = e1 . (d x e2) and the numerators of u, v, t;
// tri_exact divides and decides.
// ---------------------------------------------------------------------------
struct TriTerms {
  float det, du, dv, dt;
};
PSDR_HD TriTerms tri_terms(f3 o, f3 d, f3 p0, f3 e1, f3 e2) {
  const f3 pv = cross(d, e2);
  const f3 tv = o - p0;
  const f3 qv = cross(tv, e1);
  return TriTerms{dot(e1, pv), dot(tv, pv), dot(d, qv), dot(e2, qv)};
}
PSDR_HD bool tri_exact(const TriTerms& q, float* t, float* u, float* v) {
  const float inv = 1.0f / q.det;
  *u = q.du * inv;
  *v = q.dv * inv;
  *t = q.dt * inv;
  // every condition evaluated (no early exit)
  return (q.det != 0.f) & (*u >= 0.f) & (*u <= 1.f) & (*v >= 0.f) & (*u + *v <= 1.f) & (*t >= RAY_TMIN);
}

// ---------------------------------------------------------------------------
// Camera (src/camera.h:7-17, 33-64, 108-124): matrices are built on the host in
// fp32 and passed by value to kernels.
// ---------------------------------------------------------------------------
struct Mat4 {
  float m[4][4];
};
struct CameraK {
  Mat4 c2w, w2c, c2r, r2c;
  f3 pos;
  f3 look;    // transform_dir(camera2world, (0, 0, -1)) (src/camera.h:124)
  int H, W;
  int ortho;  // OrthographicCamera (src/camera.h:279-362)
};

PSDR_HD f3 xform_pos(const Mat4& M, f3 p) {
  float r[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = ((M.m[i][0] * p.x + M.m[i][1] * p.y) + M.m[i][2] * p.z) + M.m[i][3];
  return f3{r[0] / r[3], r[1] / r[3], r[2] / r[3]};
}
PSDR_HD f3 xform_dir(const Mat4& M, f3 d) {
  return f3{(M.m[0][0] * d.x + M.m[0][1] * d.y) + M.m[0][2] * d.z,
            (M.m[1][0] * d.x + M.m[1][1] * d.y) + M.m[1][2] * d.z,
            (M.m[2][0] * d.x + M.m[2][1] * d.y) + M.m[2][2] * d.z};
}
// Film point of a lane: lane -> pixel = lane / spp (x fastest) plus the lane's
// first two draws.
PSDR_HD f3 film_point(const CameraK& c, uint32_t lane, uint32_t spp, float u0, float u1, int* pixel) {
  const uint32_t pix = spp > 1 ? lane / spp : lane;
  const uint32_t x = pix % static_cast<uint32_t>(c.W);
  const uint32_t y = pix / static_cast<uint32_t>(c.W);
  *pixel = static_cast<int>(pix);
  return f3{u0 + static_cast<float>(x), u1 + static_cast<float>(y), 1.0f};
}
// gen_camera_ray_: perspective (src/camera.h:235-258) rays leave the camera
// position towards the film point; orthographic ones (src/camera.h:314-335)
// leave the film point's spot on the camera plane along the view direction.
PSDR_HD void camera_ray(const CameraK& c, uint32_t lane, uint32_t spp, float u0, float u1, f3* o, f3* d,
                        int* pixel) {
  const f3 p = xform_pos(c.r2c, film_point(c, lane, spp, u0, u1, pixel));
  if (c.ortho) {
    *o = xform_pos(c.c2w, f3{p.x, p.y, 0.0f});
    *d = c.look;
  } else {
    *d = normalize(xform_dir(c.c2w, p));
    *o = c.pos;
  }
}
// gen_camera_ray_with_point_ (src/camera.h:260-276, 337-361): raster pixel and
// validity of a world point; with o/d non-null also the camera ray through it.
PSDR_HD bool camera_point_ray(const CameraK& c, f3 pt, int* px, int* py, f3* o, f3* d) {
  const f3 pc = xform_pos(c.w2c, pt);
  const f3 p = xform_pos(c.c2r, pc);
  *px = static_cast<int>(floorf(p.x));
  *py = static_cast<int>(floorf(p.y));
  if (o) {
    if (c.ortho) {
      *o = xform_pos(c.c2w, f3{pc.x, pc.y, 0.0f});
      *d = c.look;
    } else {
      *o = c.pos;
      *d = normalize(pt - c.pos);
    }
  }
  return pc.z < 0.f && p.x > 0.f && p.x < static_cast<float>(c.W) && p.y > 0.f && p.y < static_cast<float>(c.H);
}
PSDR_HD bool camera_project(const CameraK& c, f3 pt, int* px, int* py) {
  return camera_point_ray(c, pt, px, py, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// DiscreteSampler::sample_reuse (src/sampler.h:98-117): binary search of the
// smallest i in [0, cnt-1] with cmf[i] >= u in floor(log2(cnt-1))+1 steps.
// ---------------------------------------------------------------------------
// The reuse of u and the pdf once entry i of cnt is picked (i the result of
// the search below).
PSDR_HD void discrete_reuse(const float* pmf, const float* cmf, int cnt, int i, float* u, float* pdf) {
  if (cnt == 1) {
    *pdf = 1.f;
    return;
  }
  const float p = pmf[i];
  *u = (*u - (i > 0 ? cmf[i - 1] : 0.f)) / p;
  *pdf = p;
}

PSDR_HD int discrete_sample(const float* pmf, const float* cmf, int cnt, float* u, float* pdf) {
  if (cnt == 1) {
    discrete_reuse(pmf, cmf, cnt, 0, u, pdf);
    return 0;
  }
  int start = 0, end = cnt - 1;
  int iters = 0;
  for (int dd = end; dd > 0; dd >>= 1) ++iters;
  for (int it = 0; it < iters; ++it) {
    const int mid = (start + end) >> 1;
    if (cmf[mid] < *u) start = min(mid + 1, end);
    else end = mid;
  }
  discrete_reuse(pmf, cmf, cnt, start, u, pdf);
  return start;
}

}  // namespace psdr
