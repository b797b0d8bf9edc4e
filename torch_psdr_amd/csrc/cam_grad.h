// Camera-parameter gradients.
//
// The reference's image derivative reaches the camera parameters only through
// ray origins (camera rays' directions feed hits whose barycentrics are
// non-differentiable, src/scene.cc:130-133):
//   * the camera vertex's direction to the viewer, wi = -normalize(p - o)
//     (Scene::hit, src/scene.cc:134-136), which the DiffuseBsdf evaluates;
//   * the primary boundary's origin term (o - detach o) . (n (x) [1, 1, -1])
//     (src/path_integrator.cc:240-244).
// Every origin is o = camera2world (x, y, 0, 1): (x, y) = (0, 0) for the
// perspective camera's position (src/camera.h:121), the raster2camera image of
// the film point for an orthographic camera ray (src/camera.h:326-330) and the
// world2camera image of the target point for an orthographic ray through a
// point (src/camera.h:343-356).  Gradient kernels add the adjoints of the
// matrix entries they use to a CAM_ACC-float block (through LDS); one thread
// then turns them into d/d(look_from, look_at, up, vfov | view_width) by
// forward-mode derivatives of the camera construction in double precision.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/psdr_hip.h"
#include "common.h"

namespace psdr {

constexpr int CAM_C = 0;   // [9] adjoint of camera2world columns 0, 1 and 3 (3 rows each)
constexpr int CAM_W = 9;   // [8] adjoint of world2camera rows 0 and 1
constexpr int CAM_R = 17;  // [8] adjoint of raster2camera rows 0 and 1
constexpr int CAM_ACC = 25;

// Adjoint g of an origin o = C (x, y, 0, 1), (x, y) = rows 0, 1 of M applied
// to (q, 1) (orthographic; perspective: (x, y) = 0).  Adds the matrix adjoints
// to acc (LDS; may be null) and returns the adjoint of q.
__device__ inline f3 cam_origin_vjp(const CameraK& c, f3 g, f3 q, const Mat4& M, int mo, float* acc) {
  if (acc) {
    atomicAdd(acc + CAM_C + 6, g.x);
    atomicAdd(acc + CAM_C + 7, g.y);
    atomicAdd(acc + CAM_C + 8, g.z);
  }
  if (!c.ortho) return splat(0.f);
  const float gx = (c.c2w.m[0][0] * g.x + c.c2w.m[1][0] * g.y) + c.c2w.m[2][0] * g.z;
  const float gy = (c.c2w.m[0][1] * g.x + c.c2w.m[1][1] * g.y) + c.c2w.m[2][1] * g.z;
  if (acc) {
    const float x = ((M.m[0][0] * q.x + M.m[0][1] * q.y) + M.m[0][2] * q.z) + M.m[0][3];
    const float y = ((M.m[1][0] * q.x + M.m[1][1] * q.y) + M.m[1][2] * q.z) + M.m[1][3];
    atomicAdd(acc + CAM_C + 0, g.x * x);
    atomicAdd(acc + CAM_C + 1, g.y * x);
    atomicAdd(acc + CAM_C + 2, g.z * x);
    atomicAdd(acc + CAM_C + 3, g.x * y);
    atomicAdd(acc + CAM_C + 4, g.y * y);
    atomicAdd(acc + CAM_C + 5, g.z * y);
    const float qq[4] = {q.x, q.y, q.z, 1.f};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      atomicAdd(acc + mo + j, gx * qq[j]);
      atomicAdd(acc + mo + 4 + j, gy * qq[j]);
    }
  }
  return f3{gx * M.m[0][0] + gy * M.m[1][0], gx * M.m[0][1] + gy * M.m[1][1], gx * M.m[0][2] + gy * M.m[1][2]};
}

__device__ inline void cam_acc_init(float* acc) {
  for (int i = threadIdx.x; i < CAM_ACC; i += blockDim.x) acc[i] = 0.f;
}
// after a barrier: one global atomic per nonzero entry
__device__ inline void cam_acc_flush(const float* acc, float* out) {
  for (int i = threadIdx.x; i < CAM_ACC; i += blockDim.x)
    if (acc[i] != 0.f) atomicAdd(out + i, acc[i]);
}

// ---- forward-mode derivatives of the camera construction (double) ----------
struct Dd {
  double v, d;
};
PSDR_HD Dd operator+(Dd a, Dd b) { return {a.v + b.v, a.d + b.d}; }
PSDR_HD Dd operator-(Dd a, Dd b) { return {a.v - b.v, a.d - b.d}; }
PSDR_HD Dd operator*(Dd a, Dd b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
PSDR_HD Dd operator/(Dd a, Dd b) { return {a.v / b.v, (a.d - (a.v / b.v) * b.d) / b.v}; }
PSDR_HD Dd dd_sqrt(Dd a) {
  const double s = sqrt(a.v);
  return {s, a.d * 0.5 / s};
}
struct V3d {
  Dd x, y, z;
};
PSDR_HD V3d operator-(const V3d& a, const V3d& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PSDR_HD V3d cross_d(const V3d& a, const V3d& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PSDR_HD V3d normalize_d(const V3d& a) {
  const Dd n = dd_sqrt((a.x * a.x + a.y * a.y) + a.z * a.z);
  return {a.x / n, a.y / n, a.z / n};
}
// inverse of a 4x4 double matrix by Gauss-Jordan elimination with partial pivoting
PSDR_HD void inverse4d(const double (&a)[4][4], double (&out)[4][4]) {
  double m[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      m[i][j] = a[i][j];
      m[i][4 + j] = i == j ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(m[r][c]) > fabs(m[p][c])) p = r;
    if (p != c)
      for (int j = 0; j < 8; ++j) {
        const double t = m[c][j];
        m[c][j] = m[p][j];
        m[p][j] = t;
      }
    const double inv = 1.0 / m[c][c];
    for (int j = 0; j < 8; ++j) m[c][j] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = m[r][c];
      for (int j = 0; j < 8; ++j) m[r][j] -= f * m[c][j];
    }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i][j] = m[i][4 + j];
}
// d(A^-1) = -A^-1 dA A^-1
PSDR_HD void dinverse4d(const double (&ai)[4][4], const double (&da)[4][4], double (&out)[4][4]) {
  double t[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += da[i][k] * ai[k][j];
      t[i][j] = s;
    }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += ai[i][k] * t[k][j];
      out[i][j] = -s;
    }
}

// camera2world C and camera2raster (screen2raster * projection) with their
// tangents for parameter k (k < 0: values only).  Parameters: look_from,
// look_at, up (look_at(), src/camera.h:33-53), then vfov or view_width.  The
// orthographic projection is ortho_mat (src/camera.h:302-312): scale
// diag(2/(r-l), 2/(t-b), 2/(n-f)) after translation (0, 0, -(n+f)/2), half
// extents hw = view_width/2, hh = H hw / W; screen2raster is
// S(W/2, -H/2, 1) T(1, -1, 0) (src/camera.h:116-118).  The perspective
// projection is not formed: no perspective origin depends on it.
PSDR_HD void camera_mats_d(const psdr_camera& cam, int k, double (&C)[4][4], double (&dC)[4][4],
                                  double (&c2r)[4][4], double (&dc2r)[4][4]) {
  V3d from, at, up;
  Dd* fp[3] = {&from.x, &from.y, &from.z};
  Dd* ap[3] = {&at.x, &at.y, &at.z};
  Dd* upp[3] = {&up.x, &up.y, &up.z};
  for (int i = 0; i < 3; ++i) {
    *fp[i] = Dd{cam.look_from[i], k == i ? 1.0 : 0.0};
    *ap[i] = Dd{cam.look_at[i], k == 3 + i ? 1.0 : 0.0};
    *upp[i] = Dd{cam.up[i], k == 6 + i ? 1.0 : 0.0};
  }
  const V3d dir = normalize_d(from - at);
  const V3d right = cross_d(normalize_d(up), dir);
  const V3d nup = cross_d(dir, right);
  const V3d cols[4] = {right, nup, dir, from};
  for (int c = 0; c < 4; ++c) {
    const Dd e[3] = {cols[c].x, cols[c].y, cols[c].z};
    for (int r = 0; r < 3; ++r) {
      C[r][c] = e[r].v;
      dC[r][c] = e[r].d;
    }
    C[3][c] = c == 3 ? 1.0 : 0.0;
    dC[3][c] = 0.0;
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) c2r[i][j] = dc2r[i][j] = 0.0;
  if (cam.kind != PSDR_CAMERA_ORTHOGRAPHIC) return;
  const Dd vw{cam.view_width, k == 9 ? 1.0 : 0.0};
  const Dd two{2.0, 0.0};
  const Dd hw = vw / two;
  const Dd hh = (Dd{static_cast<double>(cam.height), 0.0} * hw) / Dd{static_cast<double>(cam.width), 0.0};
  const Dd sx = two / (hw + hw), sy = two / (hh + hh);
  const double n = -0.01, f = -1000.0, sz = 2.0 / (n - f);
  // rows of S2R * P: x -> (W/2)(sx x + 1), y -> (-H/2)(sy y - 1)
  const double w2 = 0.5 * cam.width, h2 = -0.5 * cam.height;
  c2r[0][0] = w2 * sx.v;
  dc2r[0][0] = w2 * sx.d;
  c2r[0][3] = w2;
  c2r[1][1] = h2 * sy.v;
  dc2r[1][1] = h2 * sy.d;
  c2r[1][3] = -h2;
  c2r[2][2] = sz;
  c2r[2][3] = sz * (-(n + f) / 2.0);
  c2r[3][3] = 1.0;
}

// grad[k] += d/d(param k) of sum(acc * matrix entries): acc holds the adjoints
// of camera2world columns 0, 1, 3, of world2camera rows 0, 1 and of
// raster2camera rows 0, 1 (d(A^-1) = -A^-1 dA A^-1).
PSDR_HD void camera_param_vjp(const psdr_camera& cam, const float* acc, float* grad) {
  const bool ortho = cam.kind == PSDR_CAMERA_ORTHOGRAPHIC;
  double C[4][4], dC[4][4], c2r[4][4], dc2r[4][4], W[4][4], R[4][4], dW[4][4], dR[4][4];
  camera_mats_d(cam, -1, C, dC, c2r, dc2r);
  inverse4d(C, W);
  if (ortho) inverse4d(c2r, R);
  for (int k = 0; k < PSDR_CAMERA_GRAD; ++k) {
    camera_mats_d(cam, k, C, dC, c2r, dc2r);
    double g = 0.0;
    for (int r = 0; r < 3; ++r)
      g += acc[CAM_C + r] * dC[r][0] + acc[CAM_C + 3 + r] * dC[r][1] + acc[CAM_C + 6 + r] * dC[r][3];
    if (ortho) {
      dinverse4d(W, dC, dW);
      dinverse4d(R, dc2r, dR);
      for (int r = 0; r < 2; ++r)
        for (int j = 0; j < 4; ++j) g += acc[CAM_W + 4 * r + j] * dW[r][j] + acc[CAM_R + 4 * r + j] * dR[r][j];
    }
    grad[k] += static_cast<float>(g);
  }
}

}  // namespace psdr
