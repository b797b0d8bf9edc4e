// CPU oracle for the torch_psdr differentiable path tracer — TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, the reference's algorithm for its one hot path
// (PathIntegrator.renderC / renderD over a Scene).  It is the checker the HIP
// implementation is compared against; nothing in the product links it.  Only the
// tests, the smoke check and the benchmark's CPU-baseline leg may load it.
//
// Every routine is templated on the scalar type F:
//   F = float : the value path (renderC, and the value part of renderD);
//   F = Dual  : forward-mode derivative (one tangent direction), which restates
//               Enoki's AD semantics for renderD, including every `detach` point,
//               so that JVP(oracle) can be checked against VJP(GPU).
// Value arithmetic is ordinary IEEE fp32 (+ - * / sqrt, no FMA contraction: the
// build uses -ffp-contract=off), evaluated in the left-to-right order of the
// reference's C++ expressions.  sin/cos use the project's own Cody-Waite + Taylor
// routine (psdr_sincos below), which the GPU restates bit-for-bit.
//
// Parity pins (see tests/test_oracle_*.py):
//   PCG32 / TEA       : published pcg32 known-answer vector, TEA closed form.
//   sincos            : libm (double) within 2 ulp.
//   camera            : centre/corner rays vs closed-form look direction and fov.
//   mesh prep / edges : hand-computed tetrahedron, quad and open-strip cases.
//   closest hit       : brute force on random rays (independent BVH here).
//   LBVH              : structural invariants (leaf coverage, parent/child, key order).
//   radiance          : closed-form irradiance of a Lambertian rectangle emitter
//                       (polygon form factor) and MC convergence to it.
//   gradients         : central finite differences of this oracle's value path
//                       (texels, emission) and closed-form edge-term cases.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Scalars
// ---------------------------------------------------------------------------
struct Dual {
  float v = 0.f;  // value
  float d = 0.f;  // tangent
  Dual() = default;
  Dual(float v_) : v(v_), d(0.f) {}
  Dual(float v_, float d_) : v(v_), d(d_) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) {
  float q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }

inline float val(float x) { return x; }
inline float val(const Dual& x) { return x.v; }
inline float tan_(float) { return 0.f; }
inline float tan_(const Dual& x) { return x.d; }
inline float detach(float x) { return x; }
inline float detach(const Dual& x) { return x.v; }

inline float sqrt_(float x) { return std::sqrt(x); }
inline Dual sqrt_(Dual x) {
  float s = std::sqrt(x.v);
  return {s, x.d * 0.5f / s};
}
// enoki::safe_sqrt = sqrt(max(x, 0))
inline float safe_sqrt(float x) { return std::sqrt(std::max(x, 0.f)); }
inline Dual safe_sqrt(Dual x) { return x.v > 0.f ? sqrt_(x) : Dual{0.f, 0.f}; }
inline float abs_(float x) { return std::fabs(x); }
inline Dual abs_(Dual x) { return x.v < 0.f ? -x : x; }
// enoki::clamp: derivative passes only strictly inside the interval
inline float clamp_(float x, float lo, float hi) { return std::min(std::max(x, lo), hi); }
inline Dual clamp_(Dual x, float lo, float hi) {
  if (x.v < lo) return {lo, 0.f};
  if (x.v > hi) return {hi, 0.f};
  return x;
}
inline bool finite_(float x) { return std::isfinite(x); }

// ---------------------------------------------------------------------------
// Constants (src/types.h:114-123, src/cuda/device_programs.cu:8)
// ---------------------------------------------------------------------------
constexpr float REAL_EPS = 1e-5f;
constexpr float RAY_TMIN = 1e-3f;
constexpr float REAL_PI = 3.14159274101257324f;
constexpr float REAL_2PI = 6.28318548202514648f;
constexpr float REAL_INV_PI = 0.318309873342514038f;
constexpr float REAL_4PI = 12.5663709640502930f;     // src/types.h:120
constexpr float REAL_INV_4PI = 0.0795774683356285f;  // src/types.h:123

// ---------------------------------------------------------------------------
// sin / cos: Cody-Waite reduction by pi/2 (three-part split) and Taylor
// polynomials on [-pi/4, pi/4].  Valid for |x| < 2^10.
// ---------------------------------------------------------------------------
inline void psdr_sincos(float x, float* s, float* c) {
  const float k = std::rint(x * 0.636619746685028076f);
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
  switch (q) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
  }
}

// ---------------------------------------------------------------------------
// Small vectors
// ---------------------------------------------------------------------------
template <class F> struct V3 {
  F x, y, z;
  V3() : x(0.f), y(0.f), z(0.f) {}
  explicit V3(F a) : x(a), y(a), z(a) {}
  V3(F a, F b, F c) : x(a), y(b), z(c) {}
  F& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  const F& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class F> V3<F> operator+(const V3<F>& a, const V3<F>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class F> V3<F> operator-(const V3<F>& a, const V3<F>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class F> V3<F> operator-(const V3<F>& a) { return {-a.x, -a.y, -a.z}; }
template <class F> V3<F> operator*(const V3<F>& a, const V3<F>& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class F> V3<F> operator/(const V3<F>& a, const V3<F>& b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class F> V3<F> operator*(const V3<F>& a, F s) { return {a.x * s, a.y * s, a.z * s}; }
template <class F> V3<F> operator*(F s, const V3<F>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <class F> V3<F> operator/(const V3<F>& a, F s) { return {a.x / s, a.y / s, a.z / s}; }

template <class F> F dot(const V3<F>& a, const V3<F>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class F> V3<F> cross(const V3<F>& a, const V3<F>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class F> F squared_norm(const V3<F>& a) { return dot(a, a); }
template <class F> F norm(const V3<F>& a) { return sqrt_(squared_norm(a)); }
// normalize(v) = v * (1 / sqrt(|v|^2))
template <class F> V3<F> normalize(const V3<F>& a) {
  F inv = F(1.0f) / sqrt_(squared_norm(a));
  return a * inv;
}
template <class F> V3<float> detach(const V3<F>& a) { return {detach(a.x), detach(a.y), detach(a.z)}; }
template <class F> V3<F> lift(const V3<float>& a) { return {F(a.x), F(a.y), F(a.z)}; }
// a ray origin on scalar F: the float value, or the Dual carrying the camera tangent
template <class F> V3<F> origin_as(const V3<float>& o, const V3<Dual>& o_ad);
template <> inline V3<float> origin_as<float>(const V3<float>& o, const V3<Dual>&) { return o; }
template <> inline V3<Dual> origin_as<Dual>(const V3<float>&, const V3<Dual>& o_ad) { return o_ad; }

struct V2 {
  float x = 0.f, y = 0.f;
};

// ---------------------------------------------------------------------------
// PCG32 (O'Neill 2014, pcg32 XSH-RR 64/32) and TEA lane seeding
// (src/sampler.h:11-17, 70-80; Enoki PCG32 with the standard pcg32 constants)
// ---------------------------------------------------------------------------
constexpr uint64_t PCG32_DEFAULT_STATE = 0x853c49e6748fea9bULL;
constexpr uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;

struct Pcg32 {
  uint64_t state = 0, inc = 0;
  uint32_t next_u32() {
    const uint64_t old = state;
    state = old * PCG32_MULT + inc;
    const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
  }
  // Masked draw: the lane only advances where `active` holds (src/sampler.h:40-46).
  uint32_t next_u32(bool active) {
    const uint64_t old = state;
    if (active) state = old * PCG32_MULT + inc;
    const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
  }
  static float to_float(uint32_t u) {
    uint32_t b = (u >> 9) | 0x3f800000u;
    float f;
    std::memcpy(&f, &b, 4);
    return f - 1.0f;
  }
  float next_float() { return to_float(next_u32()); }
  float next_float(bool active) { return to_float(next_u32(active)); }
  void seed(uint64_t initstate, uint64_t initseq) {
    state = 0;
    inc = (initseq << 1) | 1u;
    next_u32();
    state += initstate;
    next_u32();
  }
};

inline uint64_t tea64(uint32_t v0, uint32_t v1) {
  uint32_t sum = 0;
  for (int i = 0; i < 4; ++i) {
    sum += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return static_cast<uint64_t>(v0) + (static_cast<uint64_t>(v1) << 32);
}

// Sampler(n) lane i: the u64 seed i + PCG32_DEFAULT_STATE is narrowed to u32
// when it enters sample_tea_64 (src/sampler.h:12-16).
inline Pcg32 lane_rng(uint32_t lane) {
  const uint32_t s = static_cast<uint32_t>(static_cast<uint64_t>(lane) + PCG32_DEFAULT_STATE);
  Pcg32 r;
  r.seed(tea64(s, lane), tea64(lane, s));
  return r;
}

// ---------------------------------------------------------------------------
// Deterministic fp32 reductions: the project fixes the summation order of
// enoki::hsum / enoki::psum as three sequential levels: chunks of 256 elements
// summed left to right; groups of 64 chunk totals scanned left to right; group
// totals scanned left to right.  A chunk's offset is (group offset + offset of
// the chunk inside its group); an element's prefix is (chunk offset + running
// sum inside the chunk).  The GPU restates the same order.
// ---------------------------------------------------------------------------
constexpr int SCAN_CHUNK = 256;
constexpr int SCAN_GROUP = 64;

inline void chunked_inclusive_scan(const float* x, int n, float* out, float* total) {
  const int nchunk = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  const int ngroup = (nchunk + SCAN_GROUP - 1) / SCAN_GROUP;
  std::vector<float> csum(nchunk, 0.f), inner(nchunk, 0.f), gsum(ngroup, 0.f), goff(ngroup, 0.f);
  for (int c = 0; c < nchunk; ++c) {
    float s = 0.f;
    for (int i = c * SCAN_CHUNK; i < std::min(n, (c + 1) * SCAN_CHUNK); ++i) s = s + x[i];
    csum[c] = s;
  }
  for (int g = 0; g < ngroup; ++g) {
    float run = 0.f;
    for (int c = g * SCAN_GROUP; c < std::min(nchunk, (g + 1) * SCAN_GROUP); ++c) {
      inner[c] = run;
      run = run + csum[c];
    }
    gsum[g] = run;
  }
  float run = 0.f;
  for (int g = 0; g < ngroup; ++g) {
    goff[g] = run;
    run = run + gsum[g];
  }
  if (total) *total = run;
  if (out) {
    for (int c = 0; c < nchunk; ++c) {
      const float off = goff[c / SCAN_GROUP] + inner[c];
      float s = 0.f;
      for (int i = c * SCAN_CHUNK; i < std::min(n, (c + 1) * SCAN_CHUNK); ++i) {
        s = s + x[i];
        out[i] = off + s;
      }
    }
  }
}
inline float chunked_sum(const float* x, int n) {
  float t = 0.f;
  chunked_inclusive_scan(x, n, nullptr, &t);
  return t;
}

// DiscreteSampler (src/sampler.h:86-124): pmf = w / sum(w), cmf = psum(pmf),
// sample_reuse binary-searches the smallest i in [0, cnt-1] with cmf[i] >= u and
// rescales u into the chosen bin.
struct Discrete {
  int cnt = 0;
  std::vector<float> pmf, cmf;
  void build(const float* w, int n) {
    cnt = n;
    pmf.assign(n, 0.f);
    cmf.assign(n, 0.f);
    if (n == 0) return;
    const float sum = chunked_sum(w, n);
    for (int i = 0; i < n; ++i) pmf[i] = w[i] / sum;
    chunked_inclusive_scan(pmf.data(), n, cmf.data(), nullptr);
  }
  // returns index, writes pdf; u is rescaled in place.
  int sample_reuse(float& u, float* pdf) const {
    if (cnt == 1) {
      *pdf = 1.f;
      return 0;
    }
    int start = 0, end = cnt - 1;
    int iters = 0;
    for (int d = end - start; d > 0; d >>= 1) ++iters;  // floor(log2(d)) + 1
    for (int it = 0; it < iters; ++it) {
      const int mid = (start + end) >> 1;
      const bool cond = cmf[mid] < u;
      if (cond) start = std::min(mid + 1, end);
      else end = mid;
    }
    const int idx = start;
    const float p = pmf[idx];
    u = (u - (idx > 0 ? cmf[idx - 1] : 0.f)) / p;
    *pdf = p;
    return idx;
  }
};

// ---------------------------------------------------------------------------
// Camera (src/camera.h:7-17, 33-64, 108-124, 235-276; near/far src/psdr.cc:27-28)
// Matrices are row-major m[r][c]; products are summed k = 0..3 left to right.
// ---------------------------------------------------------------------------
struct Mat4 {
  float m[4][4];
};
inline Mat4 mat_mul(const Mat4& a, const Mat4& b) {
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
// Inverse by cofactor expansion over 2x2 minors of the row pairs (0,1) and (2,3).
inline Mat4 mat_inverse(const Mat4& M) {
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
inline V3<float> xform_pos(const Mat4& M, const V3<float>& p) {
  float r[4];
  for (int i = 0; i < 4; ++i) r[i] = ((M.m[i][0] * p.x + M.m[i][1] * p.y) + M.m[i][2] * p.z) + M.m[i][3];
  return {r[0] / r[3], r[1] / r[3], r[2] / r[3]};
}
inline V3<float> xform_dir(const Mat4& M, const V3<float>& d) {
  return {(M.m[0][0] * d.x + M.m[0][1] * d.y) + M.m[0][2] * d.z,
          (M.m[1][0] * d.x + M.m[1][1] * d.y) + M.m[1][2] * d.z,
          (M.m[2][0] * d.x + M.m[2][1] * d.y) + M.m[2][2] * d.z};
}

// camera2world (look_at, src/camera.h:33-53: columns right, up', dir, origin)
// and, for the orthographic camera, camera2raster = screen2raster * ortho_mat
// (src/camera.h:19-31, 302-312), on any scalar (Dual: along a parameter tangent).
template <class F>
void camera_frame(const V3<F>& o, const V3<F>& t, const V3<F>& u, F C[4][4]) {
  const V3<F> dir = normalize(o - t);
  const V3<F> right = cross(normalize(u), dir);
  const V3<F> nup = cross(dir, right);
  const V3<F> cols[4] = {right, nup, dir, o};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) C[r][c] = (r < 3) ? cols[c][r] : F(c == 3 ? 1.f : 0.f);
}
template <class F>
void ortho_raster(F view_width, int H, int W, F M[4][4]) {
  const F hw = view_width / F(2.0f);
  const F hh = (F(static_cast<float>(H)) * hw) / F(static_cast<float>(W));
  const F l = F(0.f) - hw, r = hw, b = F(0.f) - hh, t = hh;
  const float n = -0.01f, f = -1000.f;
  // scale * translate, then S(W/2, -H/2, 1) T(1, -1, 0)
  const F sx = F(2.0f) / (r - l), sy = F(2.0f) / (t - b);
  const float sz = 2.0f / (n - f), tz = -(n + f) / 2.0f;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) M[i][j] = F(0.f);
  M[0][0] = F(0.5f * static_cast<float>(W)) * sx;
  M[0][3] = F(0.5f * static_cast<float>(W));
  M[1][1] = F(-0.5f * static_cast<float>(H)) * sy;
  M[1][3] = F(0.5f * static_cast<float>(H));
  M[2][2] = F(sz);
  M[2][3] = F(sz * tz);
  M[3][3] = F(1.f);
}

struct Camera {
  int H = 0, W = 0;
  bool ortho = false;
  Mat4 c2w, w2c, c2r, r2c;
  V3<float> pos;
  V3<float> look;  // transform_dir(camera2world, (0, 0, -1))
  // tangents of camera2world, world2camera and raster2camera along the
  // parameter direction of oracle_camera::d_params (d(A^-1) = -A^-1 dA A^-1)
  Mat4 dC{}, dW{}, dR{};

  void build(const float from[3], const float at[3], const float up[3], float vfov, int h, int w) {
    H = h;
    W = w;
    V3<float> o{from[0], from[1], from[2]}, t{at[0], at[1], at[2]}, u{up[0], up[1], up[2]};
    camera_frame<float>(o, t, u, c2w.m);
    w2c = mat_inverse(c2w);
    // perspective (src/camera.h:7-17), near=-0.01, far=-1000
    const float n = -0.01f, f = -1000.f;
    const float inv_nf = 1.0f / (n - f);
    float sh, ch;
    psdr_sincos(vfov * 0.5f, &sh, &ch);
    const float cot = ch / sh;
    Mat4 P{};
    std::memset(&P, 0, sizeof(P));
    P.m[0][0] = -cot;
    P.m[1][1] = -cot;
    P.m[2][2] = (n + f) * inv_nf;
    P.m[2][3] = ((-2.f * f) * n) * inv_nf;
    P.m[3][2] = 1.f;
    // screen2raster = S(W/2, -H/2, 1) * T(1, -1, 0)
    Mat4 S{}, T{};
    std::memset(&S, 0, sizeof(S));
    std::memset(&T, 0, sizeof(T));
    S.m[0][0] = 0.5f * static_cast<float>(W);
    S.m[1][1] = -0.5f * static_cast<float>(H);
    S.m[2][2] = 1.f;
    S.m[3][3] = 1.f;
    for (int i = 0; i < 4; ++i) T.m[i][i] = 1.f;
    T.m[0][3] = 1.f;
    T.m[1][3] = -1.f;
    const Mat4 s2r = mat_mul(S, T);
    c2r = mat_mul(s2r, P);
    r2c = mat_inverse(c2r);
    pos = xform_pos(c2w, V3<float>{0.f, 0.f, 0.f});
    look = xform_dir(c2w, V3<float>{0.f, 0.f, -1.f});
  }
  // OrthographicCamera (src/camera.h:279-312): the projection of ortho_mat
  void build_ortho(float view_width) {
    ortho = true;
    ortho_raster<float>(view_width, H, W, c2r.m);
    r2c = mat_inverse(c2r);
  }
  // parameter tangent (look_from, look_at, up, vfov | view_width)
  void build_tangent(const float from[3], const float at[3], const float up[3], float extent, const float dp[10]) {
    V3<Dual> o, t, u;
    for (int i = 0; i < 3; ++i) {
      o[i] = Dual(from[i], dp[i]);
      t[i] = Dual(at[i], dp[3 + i]);
      u[i] = Dual(up[i], dp[6 + i]);
    }
    Dual Cd[4][4];
    camera_frame<Dual>(o, t, u, Cd);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) dC.m[i][j] = Cd[i][j].d;
    dW = neg_sandwich(w2c, dC);
    if (ortho) {
      Dual Md[4][4];
      ortho_raster<Dual>(Dual(extent, dp[9]), H, W, Md);
      Mat4 dM;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) dM.m[i][j] = Md[i][j].d;
      dR = neg_sandwich(r2c, dM);
    }
  }
  static Mat4 neg_sandwich(const Mat4& A, const Mat4& dM) {
    Mat4 r = mat_mul(mat_mul(A, dM), A);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) r.m[i][j] = -r.m[i][j];
    return r;
  }
  // origin o = camera2world (x, y, 0, 1) with its tangent: dC (x, y, 0, 1) + C (dx, dy, 0, 0)
  V3<Dual> origin_with_tangent(const V3<float>& o, float x, float y, float dx, float dy) const {
    V3<Dual> r;
    for (int i = 0; i < 3; ++i) {
      const float t = ((dC.m[i][0] * x + dC.m[i][1] * y) + dC.m[i][3]) + (c2w.m[i][0] * dx + c2w.m[i][1] * dy);
      r[i] = Dual(o[i], t);
    }
    return r;
  }
  // gen_camera_ray_ (src/camera.h:235-258 perspective, 314-335 orthographic)
  void ray(uint32_t lane, uint32_t spp, float u0, float u1, V3<float>* o, V3<float>* d, int* px, int* py,
           V3<Dual>* o_ad = nullptr) const {
    const uint32_t pix = spp > 1 ? lane / spp : lane;
    const int x = static_cast<int>(pix % static_cast<uint32_t>(W));
    const int y = static_cast<int>(pix / static_cast<uint32_t>(W));
    const V3<float> f{u0 + static_cast<float>(x), u1 + static_cast<float>(y), 1.0f};
    const V3<float> p = xform_pos(r2c, f);
    if (ortho) {
      *o = xform_pos(c2w, V3<float>{p.x, p.y, 0.f});
      *d = look;
      if (o_ad) {
        const float dx = ((dR.m[0][0] * f.x + dR.m[0][1] * f.y) + dR.m[0][2]) + dR.m[0][3];
        const float dy = ((dR.m[1][0] * f.x + dR.m[1][1] * f.y) + dR.m[1][2]) + dR.m[1][3];
        *o_ad = origin_with_tangent(*o, p.x, p.y, dx, dy);
      }
    } else {
      *d = normalize(xform_dir(c2w, p));
      *o = pos;
      if (o_ad) *o_ad = origin_with_tangent(pos, 0.f, 0.f, 0.f, 0.f);
    }
    *px = x;
    *py = y;
  }
  // ray of gen_camera_ray_with_point_ (src/camera.h:266-275, 351-356) with
  // the origin's tangent given the point's tangent dpt
  void point_ray(const V3<float>& pt, const V3<float>& dpt, V3<float>* o, V3<float>* d, V3<Dual>* o_ad) const {
    if (!ortho) {
      *o = pos;
      *d = normalize(pt - pos);
      *o_ad = origin_with_tangent(pos, 0.f, 0.f, 0.f, 0.f);
      return;
    }
    const V3<float> pc = xform_pos(w2c, pt);
    *o = xform_pos(c2w, V3<float>{pc.x, pc.y, 0.f});
    *d = look;
    float dq[2];
    for (int r = 0; r < 2; ++r)
      dq[r] = (((dW.m[r][0] * pt.x + dW.m[r][1] * pt.y) + dW.m[r][2] * pt.z) + dW.m[r][3]) +
              ((w2c.m[r][0] * dpt.x + w2c.m[r][1] * dpt.y) + w2c.m[r][2] * dpt.z);
    *o_ad = origin_with_tangent(*o, pc.x, pc.y, dq[0], dq[1]);
  }
  // Raster area per unit world area of a small patch at x perpendicular to the
  // unit direction n (finite differences of the projection in double).  Not part
  // of the reference: it converts a boundary sample's world-measure value into
  // the raster measure of a pixel mean for the finite-difference checks.
  double raster_jacobian(const V3<float>& x, const V3<float>& n) const {
    auto proj = [&](double X, double Y, double Z, double* rx, double* ry) {
      double c[4], r[4];
      for (int i = 0; i < 4; ++i)
        c[i] = w2c.m[i][0] * X + w2c.m[i][1] * Y + w2c.m[i][2] * Z + w2c.m[i][3];
      for (int i = 0; i < 3; ++i) c[i] /= c[3];
      for (int i = 0; i < 4; ++i) r[i] = c2r.m[i][0] * c[0] + c2r.m[i][1] * c[1] + c2r.m[i][2] * c[2] + c2r.m[i][3];
      *rx = r[0] / r[3];
      *ry = r[1] / r[3];
    };
    // orthonormal a, b perpendicular to n
    double nx = n.x, ny = n.y, nz = n.z;
    const double nl = std::sqrt(nx * nx + ny * ny + nz * nz);
    nx /= nl; ny /= nl; nz /= nl;
    double ax = 0, ay = 0, az = 0;
    if (std::fabs(nx) < 0.9) ax = 1; else ay = 1;
    const double d = ax * nx + ay * ny + az * nz;
    ax -= d * nx; ay -= d * ny; az -= d * nz;
    const double al = std::sqrt(ax * ax + ay * ay + az * az);
    ax /= al; ay /= al; az /= al;
    const double bx = ny * az - nz * ay, by = nz * ax - nx * az, bz = nx * ay - ny * ax;
    const double h = 1e-4;
    double r0x, r0y, rax, ray, rbx, rby;
    proj(x.x, x.y, x.z, &r0x, &r0y);
    proj(x.x + h * ax, x.y + h * ay, x.z + h * az, &rax, &ray);
    proj(x.x + h * bx, x.y + h * by, x.z + h * bz, &rbx, &rby);
    return std::fabs((rax - r0x) * (rby - r0y) - (ray - r0y) * (rbx - r0x)) / (h * h);
  }
  // gen_camera_ray_with_point_ (src/camera.h:260-276)
  bool ray_with_point(const V3<float>& pt, int* px, int* py) const {
    const V3<float> pc = xform_pos(w2c, pt);
    const V3<float> p = xform_pos(c2r, pc);
    *px = static_cast<int>(std::floor(p.x));
    *py = static_cast<int>(std::floor(p.y));
    return pc.z < 0.f && p.x > 0.f && p.x < static_cast<float>(W) && p.y > 0.f && p.y < static_cast<float>(H);
  }
};

// ---------------------------------------------------------------------------
// Ray / triangle: Moller-Trumbore in fp32 with fixed operation order.
// u = weight of p1, v = weight of p2 (OptiX convention, device_programs.cu:45).
// Closest hit keeps the smallest t; an exactly equal t keeps the smaller global
// triangle index, so the answer does not depend on traversal order.
// ---------------------------------------------------------------------------
inline bool ray_tri(const V3<float>& o, const V3<float>& d, const V3<float>& p0, const V3<float>& e1,
                    const V3<float>& e2, float* t, float* u, float* v) {
  const V3<float> pv = cross(d, e2);
  const float det = dot(e1, pv);
  if (!(det != 0.f)) return false;
  const float inv = 1.0f / det;
  const V3<float> tv = o - p0;
  const float uu = dot(tv, pv) * inv;
  if (!(uu >= 0.f && uu <= 1.f)) return false;
  const V3<float> qv = cross(tv, e1);
  const float vv = dot(d, qv) * inv;
  if (!(vv >= 0.f && uu + vv <= 1.f)) return false;
  const float tt = dot(e2, qv) * inv;
  if (!(tt >= RAY_TMIN)) return false;
  *t = tt;
  *u = uu;
  *v = vv;
  return true;
}

struct Hit {
  int tri = -1;
  float t = std::numeric_limits<float>::infinity();
  float u = -1.f, v = -1.f;
};

// ---------------------------------------------------------------------------
// Scene data (templated on the parameter scalar)
// ---------------------------------------------------------------------------
// MAT_ENV_MAP / MAT_ENV_SH9: an environment light's bounding cube (EnvLight /
// SH9Light, src/emitter.cc:95-352), the last object.
enum MatType : int { MAT_DIFFUSE = 0, MAT_DIFFUSE_BSDF = 1, MAT_LIGHT = 2, MAT_ENV_MAP = 3, MAT_ENV_SH9 = 4 };

template <class F> struct Tri {
  V3<F> p0, p1, p2, n, t;
  F area;
  V2 uv0, uv1, uv2;
  int obj = -1;
  int vi[3] = {-1, -1, -1};
};

template <class F> struct Edge {
  V3<F> p0, p1, out, nl, nr;
  F len;
  int v0 = -1, v1 = -1;
};

struct Texture {
  int offset = 0;  // float offset into the texel array
  int h = 0, w = 0;  // 0 = constant
  int ch = 3;
};

struct Object {
  int mat = MAT_DIFFUSE;
  Texture albedo;
  Texture rough;
  Texture normal;           // DiffuseBsdf normal map (has_normal)
  bool has_normal = false;
  int light = -1;
  int tri_begin = 0, tri_count = 0;
};

template <class F> struct Light {
  V3<F> emit;
  int obj = -1;
  float power = 0.f;       // detached
  float total_area = 0.f;  // detached
  Discrete tri_sampler;
};

// Mesh preparation (src/mesh.cc:5-95).
template <class F> Tri<F> make_tri(const V3<F>& p0, const V3<F>& p1, const V3<F>& p2, V2 uv0, V2 uv1, V2 uv2) {
  Tri<F> T;
  T.p0 = p0;
  T.p1 = p1;
  T.p2 = p2;
  T.uv0 = uv0;
  T.uv1 = uv1;
  T.uv2 = uv2;
  const V3<F> a = p2 - p1;
  const V3<F> b = p0 - p2;
  const V3<F> c = p1 - p0;
  T.n = normalize(cross(c, b));
  const float duv1y = uv1.y - uv0.y;
  const float duv2y = uv2.y - uv0.y;
  T.t = normalize(F(duv2y) * c + F(duv1y) * b);
  const F al = norm(a), bl = norm(b), cl = norm(c);
  const F p = ((al + bl) + cl) / F(2.0f);
  T.area = safe_sqrt(((p * (p - al)) * (p - bl)) * (p - cl));
  return T;
}

// build_edges (src/mesh.cc:97-204) with a canonical edge order: sorted (i<j).
struct EdgeTopo {
  int i, j, l, r, kl, kr;
};
inline std::vector<EdgeTopo> edge_topology(const int* faces, int nf) {
  struct HalfEdge {
    uint64_t key;
    int face, k, slot;
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
                    trip[s][2], 3 * f + s});
    }
  }
  std::stable_sort(he.begin(), he.end(), [](const HalfEdge& x, const HalfEdge& y) { return x.key < y.key; });
  std::vector<EdgeTopo> out;
  for (size_t s = 0; s < he.size();) {
    size_t e = s;
    while (e + 1 < he.size() && he[e + 1].key == he[s].key) ++e;
    EdgeTopo t;
    t.i = static_cast<int>(he[s].key >> 32);
    t.j = static_cast<int>(he[s].key & 0xffffffffu);
    t.l = he[s].face;
    t.kl = he[s].k;
    if (e > s) {
      t.r = he[e].face;
      t.kr = he[e].k;
    } else {
      t.r = t.l;
      t.kr = t.kl;
    }
    out.push_back(t);
    s = e + 1;
  }
  return out;
}

template <class F> Edge<F> make_edge(const EdgeTopo& tp, const V3<F>* V) {
  Edge<F> E;
  E.v0 = tp.i;
  E.v1 = tp.j;
  E.p0 = V[tp.i];
  E.p1 = V[tp.j];
  const V3<F> p2l = V[tp.kl], p2r = V[tp.kr];
  V3<F> nl = normalize(cross(E.p1 - E.p0, p2l - E.p1));
  V3<F> nr = normalize(cross(p2r - E.p1, E.p1 - E.p0));
  const bool shared = tp.l != tp.r;  // `mask_unshared` in the reference is true for shared edges
  V3<F> out = shared ? normalize(nl + nr) : normalize(cross(E.p1 - E.p0, nl));
  {
    const V3<F> mid_p = E.p0 + E.p1;
    const V3<F> mid_c = p2l + p2r;
    const V3<F> cp = normalize(mid_p - mid_c);
    const bool rev = val(dot(cp, out)) < 0.f && shared && val(norm(nl - nr)) > REAL_EPS;
    if (rev) {
      nl = -nl;
      nr = -nr;
      out = -out;
    }
  }
  E.nl = nl;
  E.nr = nr;
  E.out = out;
  E.len = norm(E.p0 - E.p1);
  return E;
}

}  // namespace orc
