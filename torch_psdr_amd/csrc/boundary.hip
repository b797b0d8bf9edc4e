// Edge-sampling boundary terms of renderD (src/path_integrator.cc:11-78, 188-353).
//
// Primary (silhouette) edges, per lane of Sampler(spp_primary_edge*H*W):
//   k_prim_setup   1 draw, hypercube cell pick (10k cells), edge point, camera
//                  projection; sort key = Morton code of the pixel
//   sort           render pass only: lanes are processed in pixel-Morton order
//                  (each lane keeps its PCG stream and result; only the order
//                  changes, so neighbouring lanes of a wave trace neighbouring
//                  rays); every per-sample array from here on is indexed by
//                  the position in that order
//   k_prim_rays    the sample of each position (drawn again), rays at
//                  d -/+ 1e-5 * edge normal, queue in processing order
//   trace          both rays
//   k_prim_check   both hits valid and one of them on the edge point
//   run_depths x2  trace2light from each hit (B = 1, no light visibility),
//                  the second continuing the first one's PCG stream
//   k_prim_finish  preprocess: hypercube mass |dL / pdf|; render: the VJP of
//                  [(x_e - detach x_e) . n] dL / detach(pdf) into the edge's vertices
// Direct (shadow) edges, per lane of Sampler(spp_secondary_edge*H*W):
//   k_direct_setup 3 draws, 3-D hypercube pick (10k x 6 x 6 cells), edge and
//                  light points; queues the samples whose edge is a silhouette
//                  seen from the light point
//   k_direct       per queued sample: 3 closest hits (toward / away from the
//                  light, camera visibility), boundary_term with its
//                  hit_triangle re-parameterisation; preprocess accumulates
//                  |value|, render pushes the VJP through hit_triangle into the
//                  edge's vertices.
// The hypercube preprocessing passes reuse one Sampler(cells) stream across
// rounds, exactly as the reference does.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cam_grad.h"
#include "grad_cache.h"
#include "path.h"
#include "render.h"
#include "shading.h"
#include "traverse.h"
#include "wavefront.h"

namespace psdr {

namespace {

constexpr int BT = 256;
constexpr int BSTACK = 16;

inline int blocks(int64_t n, int t) { return static_cast<int>((n + t - 1) / t); }

struct EdgeRec {
  f3 p0, p1, out, nl, nr;
  float len;
  int v0, v1, obj;
};
__device__ __forceinline__ EdgeRec load_edge(const SceneView& sv, int e) {
  const float* g = sv.edge_geom + 16 * e;
  EdgeRec E;
  E.p0 = f3{g[0], g[1], g[2]};
  E.p1 = f3{g[3], g[4], g[5]};
  E.out = f3{g[6], g[7], g[8]};
  E.nl = f3{g[9], g[10], g[11]};
  E.nr = f3{g[12], g[13], g[14]};
  E.len = g[15];
  E.v0 = sv.edge_topo[6 * e];
  E.v1 = sv.edge_topo[6 * e + 1];
  E.obj = sv.tri_obj[sv.edge_topo[6 * e + 2]];
  return E;
}

// boundary samples aggregate their edge-vertex gradients per workgroup
constexpr int EDGE_SLOTS = 512;
using EdgeCache = LdsGradCache<EDGE_SLOTS>;

__device__ __forceinline__ void scatter_edge(const SceneView& sv, float* grad_v, const EdgeRec& E, float t, f3 g,
                                             EdgeCache* gc) {
  if (!(sv.objs[E.obj].grad_mask & PSDR_GRAD_VERTICES)) return;
  const f3 g0 = g * (1.0f - t), g1 = g * t;
  gc->add(grad_v, 0, 3 * static_cast<int64_t>(E.v0), g0);
  gc->add(grad_v, 0, 3 * static_cast<int64_t>(E.v1), g1);
}

// ---------------------------------------------------------------- primary
struct PrimRecs {
  int* edge;
  float* t;
  float* pdf_ep;
  float* pdf_cube;
  int* pix;  // -1: invalid
  float* n;  // [3][cap]
  int cap;
};
PrimRecs carve_prim(DevBuf& b, int cap) {
  PrimRecs r;
  char* p = static_cast<char*>(b.p);
  r.cap = cap;
  r.edge = reinterpret_cast<int*>(p);
  p += 4ull * cap;
  r.t = reinterpret_cast<float*>(p);
  p += 4ull * cap;
  r.pdf_ep = reinterpret_cast<float*>(p);
  p += 4ull * cap;
  r.pdf_cube = reinterpret_cast<float*>(p);
  p += 4ull * cap;
  r.pix = reinterpret_cast<int*>(p);
  p += 4ull * cap;
  r.n = reinterpret_cast<float*>(p);
  return r;
}
constexpr size_t PRIM_BYTES_PER = 4 * 5 + 12;

// PRE: preprocess round (cell = lane, state persists); else the render pass.
__device__ __forceinline__ uint32_t spread_bits16(uint32_t x) {
  x &= 0xffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

// One primary-edge sample (src/path_integrator.cc:197-214): the setup draw of
// lane i (PRE: hypercube cell i of a preprocessing round, continuing its
// state), the edge point (Scene::sample_edge_point, src/scene.cc:189-206),
// its pixel and camera ray, and the edge normal nn that offsets the side rays.
struct PrimDraw {
  uint64_t rng;  // the lane's PCG state after the draw
  int e;
  float t, pdf_ep, pdf_cube;
  bool valid;
  int px, py;
  f3 ep, o, d, nn;
};
template <bool PRE>
__device__ __forceinline__ PrimDraw prim_draw(const SceneView& sv, const CameraK& cam, int i, const uint64_t* state,
                                              int res, const float* cpmf, const float* ccmf) {
  Rng rng;
  float s, pdf_cube = 1.f;
  const float unit = 1.0f / static_cast<float>(res);
  if (PRE) {
    rng.state = state[i];
    rng.inc = lane_inc(static_cast<uint32_t>(i));
    s = (static_cast<float>(i) + rng.next()) * unit;
  } else {
    rng = lane_rng(static_cast<uint32_t>(i));
    s = rng.next();
    const int c = discrete_sample(cpmf, ccmf, res, &s, &pdf_cube);
    s = (s + static_cast<float>(c)) * unit;
  }
  PrimDraw w;
  w.rng = rng.state;
  float u = s, pdf_e;
  w.e = discrete_sample(sv.edge_pmf, sv.edge_cmf, sv.n_edges, &u, &pdf_e);
  const EdgeRec E = load_edge(sv, w.e);
  w.ep = E.p0 + u * (E.p1 - E.p0);
  w.valid = camera_point_ray(cam, w.ep, &w.px, &w.py, &w.o, &w.d);
  w.nn = cross(w.d, normalize(E.p1 - E.p0));
  w.t = u;
  w.pdf_ep = pdf_e / E.len;
  w.pdf_cube = pdf_cube;
  return w;
}
__device__ __forceinline__ void put_prim(PrimRecs& rec, int p, const PrimDraw& w, const CameraK& cam) {
  rec.edge[p] = w.e;
  rec.t[p] = w.t;
  rec.pdf_ep[p] = w.pdf_ep;
  rec.pdf_cube[p] = w.pdf_cube;
  rec.pix[p] = w.valid ? w.py * cam.W + w.px : -1;
  rec.n[p] = w.nn.x;
  rec.n[rec.cap + p] = w.nn.y;
  rec.n[2 * rec.cap + p] = w.nn.z;
}
// the side rays at d -/+ 1e-5 * nn (src/path_integrator.cc:215-218) to
// positions p and n + p; NaN directions (no traversal) for an invalid sample
__device__ __forceinline__ void put_side_rays(RaysView& rays, int n, int p, f3 o, f3 d, f3 nn, bool valid) {
  f3 da = d - nn * REAL_EPS, db = d + nn * REAL_EPS;
  if (!valid) da = db = splat(__builtin_nanf(""));
  rays.o[3 * p] = o.x; rays.o[3 * p + 1] = o.y; rays.o[3 * p + 2] = o.z;
  rays.d[3 * p] = da.x; rays.d[3 * p + 1] = da.y; rays.d[3 * p + 2] = da.z;
  const int j = n + p;
  rays.o[3 * j] = o.x; rays.o[3 * j + 1] = o.y; rays.o[3 * j + 2] = o.z;
  rays.d[3 * j] = db.x; rays.d[3 * j + 1] = db.y; rays.d[3 * j + 2] = db.z;
}

// In lane order (preprocessing rounds and the plain launch shape): the sample
// records and the post-draw PCG states of lanes [0, n).  With keys: only the
// sort keys (Morton code of the pixel, all ones when the edge point is off
// screen) and lane indices; the sample is drawn again in k_prim_rays.
template <bool PRE>
__global__ void k_prim_setup(SceneView sv, CameraK cam, int n, uint64_t* __restrict__ state, int res,
                             const float* __restrict__ cpmf, const float* __restrict__ ccmf, PrimRecs rec,
                             uint64_t* __restrict__ rng0, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PrimDraw w = prim_draw<PRE>(sv, cam, i, state, res, cpmf, ccmf);
  if (keys) {
    keys[i] = w.valid ? (spread_bits16(static_cast<uint32_t>(w.px)) | (spread_bits16(static_cast<uint32_t>(w.py)) << 1))
                      : 0xffffffffu;
    vals[i] = static_cast<uint32_t>(i);
    return;
  }
  rng0[i] = w.rng;
  put_prim(rec, i, w, cam);
}

// Position p of the processing order: its two rays go to p and n + p, its
// queue entry (lane field = p) to p.  Without order the records of lane p
// come from k_prim_setup; with order, position p draws lane order[p] itself
// and stores its record at p, so that every later per-sample kernel reads and
// writes position-indexed arrays.
__global__ void k_prim_rays(SceneView sv, CameraK cam, int n, const uint32_t* __restrict__ order, int res,
                            const float* __restrict__ cpmf, const float* __restrict__ ccmf, PrimRecs rec,
                            const uint64_t* __restrict__ rng0, RaysView rays, QueueView q0,
                            int* __restrict__ count0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) *count0 = n;
  if (p >= n) return;
  q0.lane[p] = static_cast<uint32_t>(p);
  if (order) {
    const PrimDraw w = prim_draw<false>(sv, cam, static_cast<int>(order[p]), nullptr, res, cpmf, ccmf);
    put_prim(rec, p, w, cam);
    q0.rng[p] = w.rng;
    put_side_rays(rays, n, p, w.o, w.d, w.nn, w.valid);
    return;
  }
  q0.rng[p] = rng0[p];
  // the same expressions as prim_draw, so the rays are bit-identical
  const EdgeRec E = load_edge(sv, rec.edge[p]);
  const f3 ep = E.p0 + rec.t[p] * (E.p1 - E.p0);
  int px, py;
  f3 o, d;
  (void)camera_point_ray(cam, ep, &px, &py, &o, &d);
  const f3 nn = f3{rec.n[p], rec.n[rec.cap + p], rec.n[2 * rec.cap + p]};
  put_side_rays(rays, n, p, o, d, nn, rec.pix[p] >= 0);
}

// both hits valid and one of them within 1e-4 of the edge point (src/path_integrator.cc:227-230)
__global__ void k_prim_check(SceneView sv, CameraK cam, int n, PrimRecs rec, RaysView rays) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (rec.pix[p] < 0) return;
  const EdgeRec E = load_edge(sv, rec.edge[p]);
  const f3 ep = E.p0 + rec.t[p] * (E.p1 - E.p0);
  const f3 o = f3{rays.o[3 * p], rays.o[3 * p + 1], rays.o[3 * p + 2]};
  const Its a = reconstruct(sv, rays.tri[p], rays.u[p], rays.v[p], o, true);
  const Its b = reconstruct(sv, rays.tri[n + p], rays.u[n + p], rays.v[n + p], o, true);
  const bool ok = a.valid && b.valid && (norm(a.p - ep) < 10.0f * REAL_EPS || norm(b.p - ep) < 10.0f * REAL_EPS);
  if (!ok) {
    rays.tri[p] = -1;
    rays.tri[n + p] = -1;
    rec.pix[p] = -1;
  }
}

// side B starts where side A's PCG stream ended (outputs are position-indexed)
__global__ void k_prim_requeue(int n, QueueView q0, const uint64_t* __restrict__ rng_a, int* __restrict__ count0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) *count0 = n;
  if (p >= n) return;
  q0.rng[p] = rng_a[p];
}

template <bool PRE>
__device__ __forceinline__ void prim_finish_one(const SceneView& sv, const CameraK& cam, int i, PrimRecs& rec,
                                                const float* LA, const float* LB, const uint64_t* rng_b,
                                                uint64_t* state, float* mass, const float* dLdI, uint32_t spp,
                                                float* grad_v, EdgeCache* gc, float* cacc) {
  if (PRE) state[i] = rng_b[i];
  const int pix = rec.pix[i];
  if (pix < 0) return;
  const f3 dL = f3{LA[3 * i], LA[3 * i + 1], LA[3 * i + 2]} - f3{LB[3 * i], LB[3 * i + 1], LB[3 * i + 2]};
  const float pdf_ep = rec.pdf_ep[i];
  if (PRE) {
    const f3 value = dL / pdf_ep;
    if (isfinite(sqnorm(value))) mass[i] = mass[i] + norm(value);
    return;
  }
  // value = [(x_e - detach x_e) . n] dL / detach(pdf_ep), zero in value
  const f3 value = (dL * 0.0f) / pdf_ep;
  if (!isfinite(sqnorm(value))) return;
  if (!isfinite(sqnorm(value / rec.pdf_cube[i]))) return;
  f3 g = f3{dLdI[3 * pix], dLdI[3 * pix + 1], dLdI[3 * pix + 2]};
  if (spp > 1) g = g / static_cast<float>(spp);
  g = g / rec.pdf_cube[i];
  const float gx = dot(g, dL) / pdf_ep;
  if (gx == 0.f) return;
  const f3 nn = f3{rec.n[i], rec.n[rec.cap + i], rec.n[2 * rec.cap + i]};
  const EdgeRec E = load_edge(sv, rec.edge[i]);
  f3 gep = nn * gx;
  // origin term (o - detach o) . (n (x) [1, 1, -1]); an orthographic origin is
  // the edge point's spot on the camera plane, so it also moves with the edge
  if (cacc || cam.ortho) {
    const f3 go = f3{nn.x, nn.y, -nn.z} * gx;
    const f3 ep = E.p0 + rec.t[i] * (E.p1 - E.p0);
    gep = gep + cam_origin_vjp(cam, go, ep, cam.w2c, CAM_W, cacc);
  }
  scatter_edge(sv, grad_v, E, rec.t[i], gep, gc);
}

// Position p of the processing order (records, side radiance and PCG states
// are position-indexed), so a workgroup's samples are neighbours and share
// edges.  A preprocessing round runs in lane order: position p is cell p.
template <bool PRE>
__global__ __launch_bounds__(BT) void k_prim_finish(SceneView sv, CameraK cam, int n, PrimRecs rec, const float* __restrict__ LA,
                                                    const float* __restrict__ LB, const uint64_t* __restrict__ rng_b,
                                                    uint64_t* __restrict__ state, float* __restrict__ mass,
                                                    const float* __restrict__ dLdI, uint32_t spp,
                                                    float* __restrict__ grad_v, float* __restrict__ cam_acc) {
  __shared__ int gk[PRE ? 1 : EDGE_SLOTS];
  __shared__ float gvals[PRE ? 1 : 3 * EDGE_SLOTS];
  __shared__ float cacc[PRE ? 1 : CAM_ACC];
  EdgeCache gc{gk, gvals};
  if (!PRE) {
    gc.init();
    cam_acc_init(cacc);
    __syncthreads();
  }
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n)
    prim_finish_one<PRE>(sv, cam, p, rec, LA, LB, rng_b, state, mass, dLdI, spp, grad_v,
                         &gc, cam_acc ? cacc : nullptr);
  if (!PRE) {
    __syncthreads();
    gc.flush(grad_v, grad_v);
    if (cam_acc) cam_acc_flush(cacc, cam_acc);
  }
}

// The preprocessing round of the primary boundary in one launch, one hypercube
// cell per thread: the steps of k_prim_setup<true>, k_prim_rays, the side
// rays' traversal, k_prim_check, trace2light from each side (run_depths with
// B = 1 and no light visibility) and prim_finish_one<true>, with every ray
// traced where it is made.  A cell's work is a chain of dependent traversals
// over ~10^4 cells; as separate wavefront launches (~30 of them) each level
// waits for its slowest ray and for free CUs beside the concurrent interior
// render, which made this pass the step's critical path.  Values are the
// wavefront's: the same draws in the same order, the same traversals and the
// shared path functions of path.h.
constexpr int PRE_THREADS = 64;
constexpr int PRE_STACK = 16;

// trace2light<false> (src/path_integrator.cc:80-186) from vertex `its` with
// PCG state *rng, as run_depths computes it for one queue entry: returns the
// lane radiance and leaves *rng after the path's last draw.
__device__ __forceinline__ f3 pre_path(const SceneView& sv, int D, uint32_t nL, Its its, f3 O, Rng* rng, int* stk,
                                       int* spill, int spill_stride) {
  f3 T = splat(1.f), L = splat(0.f);
  bool alive = its.valid;
  for (int d = 0; alive && d < D && !its.is_light; ++d) {
    const Its ip = its;
    // next-event estimation
    for (uint32_t i = 0; i < nL; ++i) {
      const float ru = rng->next();
      const float rv = rng->next();
      f3 dir;
      float tl;
      light_sample_ray(sv, ip.p, ru, rv, &dir, &tl);
      const HitRec h = closest_hit<PRE_STACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, ip.p, dir, stk, spill,
                                              spill_stride, vis_tmax(tl), vis_tocc(tl));
      light_sample_add(sv, ip, ru, rv, h.tri, h.u, h.v, nL, 1, 0, T, &L);
    }
    // BSDF sample with MIS: only an emitting closest hit matters
    {
      const float ru = rng->next();
      const float rv = rng->next();
      const Sampled smp = mat_sample(sv, ip, ru, rv);
      NoStats nst;
      bool skip = false;
      const HitRec seed = seek_lights(sv.seek_leaf, sv.n_seek, sv.n_seek >= 0, ip.p, smp.wo, &skip, nst);
      const HitRec h = skip ? seed
                            : closest_hit<PRE_STACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, ip.p, smp.wo, stk, spill,
                                                     spill_stride, F_INF, -F_INF, nst, &seed);
      bsdf_sample_add(sv, ip, ru, rv, h.tri, h.u, h.v, nL, 1, 0, T, &L);
    }
    // extension (none after the last depth; its draws are made all the same)
    const float ru = rng->next();
    const float rv = rng->next();
    if (d >= D - 1) break;
    const Sampled smp = mat_sample(sv, ip, ru, rv);
    const HitRec h = closest_hit<PRE_STACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, ip.p, smp.wo, stk, spill,
                                            spill_stride);
    alive = extend_path(sv, ip, ru, rv, h.tri, h.u, h.v, 0, &its, &T, &O);
  }
  if (!isfinite(sqnorm(L))) L = splat(0.f);
  return L * 1.0f;  // J0 = 1 without renderD's area jacobians
}

// k_prim_setup<true> for cell i: its draw, edge point and the side rays'
// direction; false when the edge point is not in view.
struct PreCell {
  uint64_t rng0;
  f3 ep, o, d, nn;
  float pdf_ep;
};
__device__ __forceinline__ bool pre_cell(const SceneView& sv, const CameraK& cam, int n, int i,
                                         const uint64_t* __restrict__ state, PreCell* c) {
  const PrimDraw w = prim_draw<true>(sv, cam, i, state, n, nullptr, nullptr);
  c->rng0 = w.rng;
  c->ep = w.ep;
  c->o = w.o;
  c->d = w.d;
  c->nn = w.nn;
  c->pdf_ep = w.pdf_ep;
  return w.valid;
}

// k_prim_check: both side rays hit and one of them lies on the edge point
__device__ __forceinline__ bool pre_check(const SceneView& sv, const PreCell& c, const HitRec& a, const HitRec& b) {
  const Its ia = reconstruct(sv, a.tri, a.u, a.v, c.o, true);
  const Its ib = reconstruct(sv, b.tri, b.u, b.v, c.o, true);
  return ia.valid && ib.valid && (norm(ia.p - c.ep) < 10.0f * REAL_EPS || norm(ib.p - c.ep) < 10.0f * REAL_EPS);
}

// One cell per thread (any number of light samples per vertex).
__global__ __launch_bounds__(PRE_THREADS) void k_prim_pre(SceneView sv, CameraK cam, int n, int D, uint32_t nL,
                                                           uint64_t* __restrict__ state, float* __restrict__ mass,
                                                           int* __restrict__ spill) {
  // the primary boundary's chain is the step's critical path
  __builtin_amdgcn_s_setprio(3);
  __shared__ int stk[PRE_STACK * PRE_THREADS];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  if (i >= n) return;
  int* sp = spill + i;
  PreCell c;
  bool ok = pre_cell(sv, cam, n, i, state, &c);
  HitRec ha{-1, F_INF, -1.f, -1.f}, hb = ha;
  if (ok) {
    ha = closest_hit<PRE_STACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, c.o, c.d - c.nn * REAL_EPS, stk, sp, stride);
    hb = closest_hit<PRE_STACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, c.o, c.d + c.nn * REAL_EPS, stk, sp, stride);
    ok = pre_check(sv, c, ha, hb);
  }
  if (!ok) {
    // the paths of an invalid sample end at once: no draws, no radiance
    state[i] = c.rng0;
    return;
  }
  // both sides' paths from the camera-side vertex (origin: the camera, or the
  // orthographic ray's own origin), side B continuing side A's PCG stream
  const f3 O = cam.ortho ? c.o : cam.pos;
  Rng r{c.rng0, lane_inc(static_cast<uint32_t>(i))};
  const f3 LA = pre_path(sv, D, nL, reconstruct(sv, ha.tri, ha.u, ha.v, O, true), O, &r, stk, sp, stride);
  const f3 LB = pre_path(sv, D, nL, reconstruct(sv, hb.tri, hb.u, hb.v, O, true), O, &r, stk, sp, stride);
  // prim_finish_one<true>
  state[i] = r.state;
  const f3 value = (LA - LB) / c.pdf_ep;
  if (isfinite(sqnorm(value))) mass[i] = mass[i] + norm(value);
}

// One light sample per vertex: four lanes per cell.  A round traces the
// cell's rays of one depth on separate lanes (the two side rays; then the
// light-sample, BSDF-sample and extension rays of each depth of side A, then
// of side B); lane 0 of the group draws, shades and makes the next round's
// rays.  A cell's latency is then one traversal per depth instead of three.
constexpr int PRE_GROUP = 4;
struct PreRay {
  f3 o, d;
  float tmax, tocc;
  HitRec h;  // in: the seed (tri -1: none); out: the hit
  int on;    // trace this slot in the round
};

__global__ __launch_bounds__(PRE_THREADS) void k_prim_pre_group(SceneView sv, CameraK cam, int n, int D,
                                                                 uint64_t* __restrict__ state,
                                                                 float* __restrict__ mass, int* __restrict__ spill) {
  __builtin_amdgcn_s_setprio(3);
  __shared__ int stk[PRE_STACK * PRE_THREADS];
  __shared__ PreRay slot[PRE_THREADS];
  const int lane = threadIdx.x, sub = lane % PRE_GROUP;
  const int i = (blockIdx.x * blockDim.x + lane) / PRE_GROUP;  // the cell
  const int gtid = blockIdx.x * blockDim.x + lane;
  const int stride = gridDim.x * blockDim.x;
  PreRay* my = &slot[lane - sub];  // the group's slots
  const bool lead = sub == 0 && i < n;
  // lane 0's cell state
  PreCell c;
  Rng rng{0, 0};
  f3 T = splat(1.f), L = splat(0.f), LA = splat(0.f), O = splat(0.f);
  Its its;
  int phase = 0;  // 0 side rays, 1 side A, 2 side B, 3 done
  int d = 0;
  float ru[3] = {0.f, 0.f, 0.f}, rv[3] = {0.f, 0.f, 0.f};
  slot[lane].on = 0;
  __syncthreads();
  if (lead) {
    const bool ok = pre_cell(sv, cam, n, i, state, &c);
    rng = Rng{c.rng0, lane_inc(static_cast<uint32_t>(i))};
    if (ok) {
      my[0] = PreRay{c.o, c.d - c.nn * REAL_EPS, F_INF, -F_INF, HitRec{-1, F_INF, -1.f, -1.f}, 1};
      my[1] = PreRay{c.o, c.d + c.nn * REAL_EPS, F_INF, -F_INF, HitRec{-1, F_INF, -1.f, -1.f}, 1};
    } else {
      state[i] = c.rng0;  // an invalid sample's paths end at once
      phase = 3;
    }
  }
  NoStats nst;
  const f3 nan3 = splat(__builtin_nanf(""));
  for (;;) {
    __syncthreads();
    // trace this lane's slot
    PreRay& s = slot[lane];
    if (s.on) {
      const HitRec seed = s.h;
      s.h = closest_hit<PRE_STACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, s.o, s.d, stk, spill + gtid, stride, s.tmax,
                                   s.tocc, nst, &seed);
    }
    __syncthreads();
    bool more = false;
    if (lead && phase < 3) {
      if (phase == 0) {
        const HitRec ha = my[0].h, hb = my[1].h;
        if (!pre_check(sv, c, ha, hb)) {
          state[i] = c.rng0;
          phase = 3;
        } else {
          O = cam.ortho ? c.o : cam.pos;
          its = reconstruct(sv, ha.tri, ha.u, ha.v, O, true);
          phase = 1;
          d = 0;
          my[3].h = hb;  // side B's first hit, kept for its path (slot 3 is never traced)
        }
      } else {
        // resolve the depth traced in this round
        const Its ip = its;
        light_sample_add(sv, ip, ru[0], rv[0], my[0].h.tri, my[0].h.u, my[0].h.v, 1, 1, 0, T, &L);
        bsdf_sample_add(sv, ip, ru[1], rv[1], my[1].h.tri, my[1].h.u, my[1].h.v, 1, 1, 0, T, &L);
        bool alive = false;
        if (d < D - 1) alive = extend_path(sv, ip, ru[2], rv[2], my[2].h.tri, my[2].h.u, my[2].h.v, 0, &its, &T, &O);
        ++d;
        if (!alive) its.valid = false;
      }
      // paths that end here: side A hands over to side B, side B finishes
      while (phase == 1 || phase == 2) {
        if (its.valid && d < D && !its.is_light) break;
        if (!isfinite(sqnorm(L))) L = splat(0.f);
        L = L * 1.0f;
        if (phase == 1) {
          LA = L;
          L = splat(0.f);
          T = splat(1.f);
          O = cam.ortho ? c.o : cam.pos;
          its = reconstruct(sv, my[3].h.tri, my[3].h.u, my[3].h.v, O, true);
          d = 0;
          phase = 2;
        } else {
          state[i] = rng.state;
          const f3 value = (LA - L) / c.pdf_ep;
          if (isfinite(sqnorm(value))) mass[i] = mass[i] + norm(value);
          phase = 3;
        }
      }
      my[0].on = my[1].on = my[2].on = 0;
      if (phase == 1 || phase == 2) {
        // the next depth's rays from `its`
        for (int k = 0; k < 3; ++k) {
          ru[k] = rng.next();
          rv[k] = rng.next();
        }
        float tl;
        f3 dir;
        light_sample_ray(sv, its.p, ru[0], rv[0], &dir, &tl);
        my[0] = PreRay{its.p, dir, vis_tmax(tl), vis_tocc(tl), HitRec{-1, F_INF, -1.f, -1.f}, 1};
        const Sampled sb = mat_sample(sv, its, ru[1], rv[1]);
        bool skip = false;
        const HitRec seed = seek_lights(sv.seek_leaf, sv.n_seek, sv.n_seek >= 0, its.p, sb.wo, &skip, nst);
        my[1] = PreRay{its.p, sb.wo, F_INF, -F_INF, seed, skip ? 0 : 1};
        const bool ext = d < D - 1;
        my[2] = PreRay{its.p, ext ? mat_sample(sv, its, ru[2], rv[2]).wo : nan3, F_INF, -F_INF,
                       HitRec{-1, F_INF, -1.f, -1.f}, ext ? 1 : 0};
        more = true;
      }
    }
    if (!__syncthreads_or(more)) break;
  }
}

// ---------------------------------------------------------------- direct
__device__ __forceinline__ f3 vjp_norm(f3 x, f3 gy) {
  const float r = 1.0f / sqrtf(sqnorm(x));
  const f3 y = x * r;
  return (gy - y * dot(y, gy)) * r;
}

struct DirectOut {
  bool valid;
  int pix;
  f3 value;  // preprocess value (non-differentiable form)
};

// The first steps of eval_direct_boundary (src/path_integrator.cc:287-299):
// edge point, light point, and whether the edge is a silhouette seen from the
// light point.  Most samples of the preprocessing pass end here, before any
// traversal, so a first kernel runs only this part and queues the samples that
// pass for the traversing kernel, which repeats it (bit for bit).
struct DirectSetup {
  bool valid;
  int e;
  float ue, pdf_ep, pdf_light;
  EdgeRec E;
  f3 ep, dir;
  LightPoint lp;
};

__device__ __forceinline__ DirectSetup direct_setup(const SceneView& sv, float sx, float sy, float sz) {
  DirectSetup d;
  d.valid = false;
  if (sv.n_edges == 0) return d;
  d.ue = sx;
  float pdf_e;
  d.e = discrete_sample(sv.edge_pmf, sv.edge_cmf, sv.n_edges, &d.ue, &pdf_e);
  d.E = load_edge(sv, d.e);
  d.ep = d.E.p0 + d.ue * (d.E.p1 - d.E.p0);
  d.pdf_ep = pdf_e / d.E.len;
  float y = sy, pdf_choose;
  const int obj = discrete_sample(sv.obj_pmf, sv.obj_cmf, sv.n_objects, &y, &pdf_choose);
  const int l = sv.objs[obj].light;
  if (l < 0) return d;
  d.lp = sample_light_point(sv, l, y, sz);
  d.pdf_light = d.lp.pdf * pdf_choose;
  d.dir = normalize(d.ep - d.lp.p);
  d.valid = (dot(d.E.nl, d.dir) * dot(d.E.nr, d.dir) < 0.f) && norm(d.E.nl - d.E.nr) > REAL_EPS;
  return d;
}

// eval_direct_boundary (src/path_integrator.cc:287-329) + boundary_term
// (src/path_integrator.cc:31-78).  Render pass (!PRE): the adjoint of the value
// is dLdI[pixel] * gscale and is pushed into the edge's vertices.
template <bool PRE>
__device__ DirectOut eval_direct(const SceneView& sv, const CameraK& cam, float sx, float sy, float sz, int* stk,
                                 int* spill, int spill_stride, float gscale, float* rec) {
  DirectOut r;
  r.valid = false;
  r.pix = -1;
  r.value = splat(0.f);
  const DirectSetup ds = direct_setup(sv, sx, sy, sz);
  if (!ds.valid) return r;
  const int e = ds.e;
  const float ue = ds.ue, pdf_ep = ds.pdf_ep, pdf_light = ds.pdf_light;
  const EdgeRec& E = ds.E;
  const f3 ep = ds.ep, dir = ds.dir;
  const LightPoint& lp = ds.lp;
  bool valid;
  const float tl = norm(ep - lp.p);
  HitRec h = closest_hit<BSTACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, ep, -dir, stk, spill, spill_stride,
                                 vis_tmax(tl), vis_tocc(tl));
  const Its s0 = reconstruct(sv, h.tri, h.u, h.v, ep, true);
  valid = s0.valid && s0.is_light && sqnorm(s0.p - lp.p) < REAL_EPS;
  if (!valid) return r;
  h = closest_hit<BSTACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, ep, dir, stk, spill, spill_stride);
  const Its d0 = reconstruct(sv, h.tri, h.u, h.v, ep, true);
  valid = d0.valid && !d0.is_light;
  if (!valid) return r;
  int px, py;
  f3 co, cd;
  if (!camera_point_ray(cam, d0.p, &px, &py, &co, &cd)) return r;
  const float tc = norm(d0.p - co);
  h = closest_hit<BSTACK>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, co, cd, stk, spill, spill_stride,
                          vis_tmax(tc), vis_tocc(tc));
  const Its i2 = reconstruct(sv, h.tri, h.u, h.v, co, true);
  if (!(i2.valid && !i2.is_light && sqnorm(i2.p - d0.p) < REAL_EPS)) return r;
  const float pdf_xw = ((s0.d2 * pdf_ep) * pdf_light) / dot(s0.n, dir);
  // boundary_term
  const f3 xs = s0.p, xd = d0.p, xb = ep, ns = s0.n, nd = d0.n;
  const f3 db = normalize(xb - xs);
  f3 ev = normalize(E.p1 - E.p0);
  if (dot(cross(db, ev), E.out) < 0.f) ev = -ev;
  const f3 n_plane = cross(db, ev);
  const f3 edge_proj = cross(nd, n_plane);
  const float cos_s = fabsf(dot(ns, db));
  const float cos_b = dot(ev, db);
  const float cos_d = dot(edge_proj, db);
  const float sin_b = safe_sqrt(1.0f - cos_b * cos_b);
  const float sin_d = safe_sqrt(1.0f - cos_d * cos_d);
  const float len_sb = norm(xb - xs);
  const float len_sd = norm(xd - xs);
  const float Jb = ((len_sd * len_sb) * sin_b) / (sin_d * cos_s);
  const float dG = (cos_s * dot(nd, -db)) / (len_sd * len_sd);
  const int ls = s0.light;
  const f3 Ts = light_radiance(sv, ls, s0);
  const f3 Td = mat_eval(sv, d0, -cd);
  r.pix = py * cam.W + px;
  if (PRE) {
    r.value = ((splat(dG) * Jb) / pdf_xw * Ts) * Td;
    r.valid = isfinite(sqnorm(r.value));
    return r;
  }
  const f3 nep = cross(nd, edge_proj);
  const f3 fbk = nep * dG;  // value = (x_hit - detach x_hit) * fbk * Jb / pdf_xw * Ts * Td
  const f3 v0 = (((fbk * 0.0f) * Jb) / pdf_xw * Ts) * Td;
  r.valid = isfinite(sqnorm(v0));
  if (!r.valid) return r;
  // The vertex adjoint is linear in the pixel adjoint g: gxb = M g.  Column c of
  // M is the VJP below for g = e_c (gscale folded in), so the pass that knows
  // dL/dI only forms M g and scatters it.
  // hit_triangle(xs, db, xd, edge_proj, nep) (src/path_integrator.cc:11-29)
  const f3 E1 = edge_proj, E2 = nep;
  const f3 T = xs - xd;
  const f3 P = cross(db, E2);
  const f3 Q = cross(T, E1);
  const float pe = dot(P, E1);
  const float cl = fminf(fmaxf(pe, REAL_EPS), 1.0f - REAL_EPS);
  const float inv = 1.0f / cl;
  const bool unclamped = !(pe < REAL_EPS) && !(pe > 1.0f - REAL_EPS);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const f3 g = f3{c == 0 ? gscale : 0.f, c == 1 ? gscale : 0.f, c == 2 ? gscale : 0.f};
    const f3 K = (((g * fbk) * Jb) / pdf_xw * Ts) * Td;  // adjoint of x_hit
    const float gu = dot(K, E1), gv = dot(K, E2);
    f3 gP = T * (gu * inv);
    const float ginv = gu * dot(P, T) + gv * dot(Q, db);
    f3 gD = Q * (gv * inv);
    if (unclamped) gP = gP + E1 * (-ginv * inv * inv);
    gD = gD + cross(E2, gP);
    const f3 gxb = vjp_norm(xb - xs, gD);
    rec[3 * c] = gxb.x;
    rec[3 * c + 1] = gxb.y;
    rec[3 * c + 2] = gxb.z;
  }
  rec[9] = __int_as_float(r.pix);
  rec[10] = __int_as_float(e);
  rec[11] = ue;
  return r;
}

// Samples of a direct-boundary pass: PRE, one per cell (accumulates |value|
// into mass); render pass, position p of the processing order draws lane
// order[p] and owns the 12-float record at p (the 3x3 map from the pixel
// adjoint to the edge-point adjoint, pixel, edge, edge parameter; pixel -1 when
// the sample is invalid).  k_direct_setup draws each sample and queues those
// whose edge is a silhouette from the light point (p, draws, gradient scale),
// or every sample when compact is 0 (the reference's launch shape);
// k_direct traces the queued ones.
template <bool PRE>
__global__ __launch_bounds__(BT) void k_direct_setup(SceneView sv, int n, uint64_t* __restrict__ state, int r0,
                                                     int r1, int r2, const float* __restrict__ cpmf,
                                                     const float* __restrict__ ccmf, uint32_t spp,
                                                     float* __restrict__ drec, const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ sorted_cells,
                                                     int* __restrict__ qidx, float4* __restrict__ qval,
                                                     int* __restrict__ qcount, int compact) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  bool want = false;
  float s[3] = {0.f, 0.f, 0.f};
  float gscale = 0.f;
  if (p < n) {
    const int cells = r0 * r1 * r2;
    const float u0 = 1.0f / static_cast<float>(r0), u1 = 1.0f / static_cast<float>(r1),
                u2 = 1.0f / static_cast<float>(r2);
    const int i = order ? static_cast<int>(order[p]) : p;
    Rng rng;
    if (PRE) {
      rng.state = state[i];
      rng.inc = lane_inc(static_cast<uint32_t>(i));
      for (int k = 0; k < 3; ++k) s[k] = rng.next();
      state[i] = rng.state;
      const int c0 = i / (r1 * r2), c1 = (i / r2) % r1, c2 = i % r2;
      s[0] = (static_cast<float>(c0) + s[0]) * u0;
      s[1] = (static_cast<float>(c1) + s[1]) * u1;
      s[2] = (static_cast<float>(c2) + s[2]) * u2;
    } else {
      rng = lane_rng(static_cast<uint32_t>(i));
      for (int k = 0; k < 3; ++k) s[k] = rng.next();
      float pdf_cube = 1.f;
      // the cell k_direct_keys picked for this lane (the sort key), else the search
      int c;
      if (sorted_cells) {
        c = static_cast<int>(sorted_cells[p]);
        discrete_reuse(cpmf, ccmf, cells, c, &s[2], &pdf_cube);
      } else {
        c = discrete_sample(cpmf, ccmf, cells, &s[2], &pdf_cube);
      }
      const int c0 = c / (r1 * r2), c1 = (c / r2) % r1, c2 = c % r2;
      s[0] = (s[0] + static_cast<float>(c0)) * u0;
      s[1] = (s[1] + static_cast<float>(c1)) * u1;
      s[2] = (s[2] + static_cast<float>(c2)) * u2;
      gscale = 1.0f / pdf_cube;
      if (spp > 1) gscale = gscale / static_cast<float>(spp);
      drec[12ll * p + 9] = __int_as_float(-1);
    }
    want = !compact || direct_setup(sv, s[0], s[1], s[2]).valid;
  }
  __shared__ int alloc_lds[BT / 64 + 1];
  const int q = block_alloc(want, qcount, alloc_lds);
  if (want) {
    qidx[q] = p;
    qval[q] = make_float4(s[0], s[1], s[2], gscale);
  }
}

template <bool PRE>
__global__ __launch_bounds__(BT) void k_direct(SceneView sv, CameraK cam, const int* __restrict__ qcount,
                                               const int* __restrict__ qidx, const float4* __restrict__ qval,
                                               float* __restrict__ mass, float* __restrict__ drec,
                                               int* __restrict__ spill) {
  __shared__ int stk[BSTACK * BT];
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  const int nq = *qcount;
  for (int q = gtid; q < nq; q += stride) {
    const int p = qidx[q];
    const float4 v = qval[q];
    if (PRE) {
      DirectOut o = eval_direct<true>(sv, cam, v.x, v.y, v.z, stk, spill + gtid, stride, 0.f, nullptr);
      if (o.valid) {
        f3 w = o.value;
        w.x = w.x < REAL_EPS ? REAL_EPS : w.x;
        w.y = w.y < REAL_EPS ? REAL_EPS : w.y;
        w.z = w.z < REAL_EPS ? REAL_EPS : w.z;
        mass[p] = mass[p] + norm(w);
      }
    } else {
      eval_direct<false>(sv, cam, v.x, v.y, v.z, stk, spill + gtid, stride, v.w, drec + 12ll * p);
    }
  }
}

// Direct-boundary gradients from the records: g = dL/dI at the sample's pixel,
// edge-point adjoint M g, split onto the edge's two vertices (aggregated per
// workgroup; positions are in cell order, so neighbours share edges).
__global__ __launch_bounds__(BT) void k_direct_scatter(SceneView sv, int n, const float* __restrict__ drec,
                                                       const float* __restrict__ dLdI, float* __restrict__ grad_v) {
  __shared__ int gk[EDGE_SLOTS];
  __shared__ float gvals[3 * EDGE_SLOTS];
  EdgeCache gc{gk, gvals};
  gc.init();
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    const float* rec = drec + 12ll * p;
    const int pix = __float_as_int(rec[9]);
    if (pix >= 0) {
      const f3 g = f3{dLdI[3 * pix], dLdI[3 * pix + 1], dLdI[3 * pix + 2]};
      const f3 gxb = (f3{rec[0], rec[1], rec[2]} * g.x + f3{rec[3], rec[4], rec[5]} * g.y) +
                     f3{rec[6], rec[7], rec[8]} * g.z;
      if (gxb.x != 0.f || gxb.y != 0.f || gxb.z != 0.f) {
        const EdgeRec E = load_edge(sv, __float_as_int(rec[10]));
        scatter_edge(sv, grad_v, E, rec[11], gxb, &gc);
      }
    }
  }
  __syncthreads();
  gc.flush(grad_v, grad_v);
}

// Sort key of a direct-boundary lane: its hypercube cell (edge position major),
// drawn exactly as k_direct<false> draws it.
__global__ void k_direct_keys(int n, int cells, const float* __restrict__ cpmf, const float* __restrict__ ccmf,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rng rng = lane_rng(static_cast<uint32_t>(i));
  rng.next();
  rng.next();
  float s2 = rng.next(), pdf;
  keys[i] = static_cast<uint32_t>(discrete_sample(cpmf, ccmf, cells, &s2, &pdf));
  vals[i] = static_cast<uint32_t>(i);
}

__global__ void k_fill_f32(float* p, int n, float v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_zero_count(int* c) { *c = 0; }

__global__ void k_seed_states(uint64_t* st, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) st[i] = lane_rng(static_cast<uint32_t>(i)).state;
}

}  // namespace

int fill_f32(float* p, int n, float v, hipStream_t st) {
  if (n <= 0) return PSDR_OK;
  hipLaunchKernelGGL(k_fill_f32, dim3((n + BT - 1) / BT), dim3(BT), 0, st, p, n, v);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

namespace {

int g_cus_b = 0;
int cus() {
  if (!g_cus_b) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    g_cus_b = (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) ? p.multiProcessorCount
                                                                                         : 256;
  }
  return g_cus_b;
}

}  // namespace

// One primary pass over n lanes up to both sides' radiance (PRE: a preprocess
// round over n = cells, which also accumulates the cell masses).
static int primary_pass(psdr_scene* s, const psdr_config& cfg, const CameraK& cam, bool pre, int n, int res,
                        const float* cpmf, const float* ccmf, float* mass, hipStream_t st) {
  RenderSlot& slot = *s->ws->cur;
  const SceneView sv = s->view();
  const uint32_t nL = cfg.mis_light_samples;
  const int Kmax = static_cast<int>(nL + 1 + 1);
  const int64_t ray_cap = static_cast<int64_t>(n) * Kmax;
  if (slot.b_rec.ensure(PRIM_BYTES_PER * n) || slot.b_rays.ensure(RAY_BYTES_PER * 2 * n) || slot.b_hits.ensure(12ull * 2 * n) ||
      slot.b_q0.ensure(QUEUE_BYTES_PER * n) || slot.b_LA.ensure(12ull * n) || slot.b_LB.ensure(12ull * n) ||
      slot.b_rngA.ensure(8ull * n) || slot.b_rngB.ensure(8ull * n) || slot.p_qa.ensure(QUEUE_BYTES_PER * n) ||
      slot.p_qb.ensure(QUEUE_BYTES_PER * n) || slot.p_rays.ensure(RAY_BYTES_PER * ray_cap) ||
      slot.p_hits.ensure(12ull * ray_cap) || slot.p_rays2.ensure(RAY_BYTES_PER * ray_cap) ||
      slot.p_hits2.ensure(12ull * ray_cap) || slot.b_counters.ensure(4 * 256))
    return grow_failed();
  PrimRecs rec = carve_prim(slot.b_rec, n);
  RaysView r0 = carve_rays(slot.b_rays, slot.b_hits, 2ll * n);
  const QueueView q0 = carve_queue(slot.b_q0, n);
  const QueueView qs[2] = {carve_queue(slot.p_qa, n), carve_queue(slot.p_qb, n)};
  const RaysView rs[2] = {carve_rays(slot.p_rays, slot.p_hits, ray_cap), carve_rays(slot.p_rays2, slot.p_hits2, ray_cap)};
  int* cnt = slot.b_counters.as<int>();
  int rc = zero_ints(cnt, 256, st);
  if (rc) return rc;
  uint64_t* state = slot.b_state.as<uint64_t>();
  // the lane's PCG state after the setup draw; pass B's output overwrites it later
  uint64_t* rng0 = slot.b_rngB.as<uint64_t>();
  // render pass: process lanes in pixel-Morton order (the preprocess lanes are
  // hypercube cells, already in edge order)
  const uint32_t* order = nullptr;
  uint32_t *k0 = nullptr, *v0 = nullptr;
  const bool sorted = !pre && !s->plain;
  if (sorted) {
    if (slot.b_keys.ensure(8ull * n) || slot.b_vals.ensure(8ull * n) || slot.b_hist.ensure(4 * radix_hist_words(n)))
      return grow_failed();
    k0 = slot.b_keys.as<uint32_t>();
    v0 = slot.b_vals.as<uint32_t>();
  }
  if (pre)
    hipLaunchKernelGGL(k_prim_setup<true>, dim3(blocks(n, BT)), dim3(BT), 0, st, sv, cam, n, state, res, cpmf, ccmf,
                       rec, rng0, nullptr, nullptr);
  else
    hipLaunchKernelGGL(k_prim_setup<false>, dim3(blocks(n, BT)), dim3(BT), 0, st, sv, cam, n, state, res, cpmf, ccmf,
                       rec, rng0, k0, v0);
  PSDR_HIP_CHECK(hipGetLastError());
  if (sorted) {
    // keys are 2 x 16-bit Morton codes; a 2 x 11-bit code needs 3 byte passes
    const int passes = (cam.W <= 2048 && cam.H <= 2048) ? 3 : 4;
    rc = radix_sort_pairs(k0, v0, k0 + n, v0 + n, slot.b_hist.as<uint32_t>(), n, passes, st);
    if (rc) return rc;
    order = (passes & 1) ? v0 + n : v0;
  }
  hipLaunchKernelGGL(k_prim_rays, dim3(blocks(n, BT)), dim3(BT), 0, st, sv, cam, n, order, res, cpmf, ccmf, rec, rng0,
                     r0, q0, cnt);
  PSDR_HIP_CHECK(hipGetLastError());
  const int tgrid = trace_grid(2ll * n);
  if (slot.p_spill.ensure(std::max(slot.p_spill.bytes, spill_bytes(tgrid)))) return grow_failed();
  // the primary term's chain (preprocess, sort, both sides' paths) is the
  // longest of a renderD step: its traversal and shading waves take issue
  // priority over the concurrent interior and direct-boundary kernels
  const int prio = s->plain ? 0 : 1;
  rc = trace_rays(sv, r0.o, r0.d, 2 * n, nullptr, 1, 0, nullptr, 0, r0.tri, nullptr, r0.u, r0.v, slot.p_spill.as<int>(),
                  tgrid, st, 0, prio);
  if (rc) return rc;
  if (s->plain) PSDR_HIP_CHECK(hipStreamSynchronize(st));
  hipLaunchKernelGGL(k_prim_check, dim3(blocks(n, BT)), dim3(BT), 0, st, sv, cam, n, rec, r0);
  WaveArgs wa;
  wa.cam = cam;
  wa.D = static_cast<int>(cfg.max_bounce);
  wa.nL = nL;
  wa.nB = 1;
  wa.elv = 0;
  wa.ad_jac = 0;
  wa.keep = 0;
  wa.cache = CacheView{nullptr, nullptr, nullptr, n};
  wa.lane_L = slot.b_LA.as<float>();
  wa.lane_rng_out = slot.b_rngA.as<uint64_t>();
  wa.counters = cnt;
  wa.plain = s->plain ? 1 : 0;
  wa.prio = prio;
  wa.spill = &slot.p_spill;
  wa.lane_of = order;  // queue lanes are positions; the PCG streams those of order[p]
  rc = run_depths(s, wa, q0, r0, qs, rs, n, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_prim_requeue, dim3(blocks(n, BT)), dim3(BT), 0, st, n, q0, slot.b_rngA.as<uint64_t>(),
                     cnt + 128);
  wa.lane_L = slot.b_LB.as<float>();
  wa.lane_rng_out = slot.b_rngB.as<uint64_t>();
  wa.counters = cnt + 128;
  rc = run_depths(s, wa, q0, offset_rays(r0, n), qs, rs, n, st);
  if (rc) return rc;
  if (pre) {
    hipLaunchKernelGGL(k_prim_finish<true>, dim3(blocks(n, BT)), dim3(BT), 0, st, sv, cam, n, rec,
                       slot.b_LA.as<float>(), slot.b_LB.as<float>(), slot.b_rngB.as<uint64_t>(), state, mass, nullptr,
                       cfg.spp_primary_edge, nullptr, nullptr);
    PSDR_HIP_CHECK(hipGetLastError());
  }
  return PSDR_OK;
}

// Number of primary / direct render-pass lanes (spp_edge * H * W), or -1 (error set).
static int edge_lanes(uint32_t spp, const CameraK& cam, const char* what) {
  const int64_t n64 = static_cast<int64_t>(spp) * cam.H * cam.W;
  if (n64 > 0x7fffffff / 8) {
    set_error(PSDR_ERR_ARG, std::string(what) + ": too many samples");
    return -1;
  }
  return static_cast<int>(n64);
}

int boundary_primary_prepare(psdr_scene* s, const psdr_config& cfg, const psdr_camera& camd, hipStream_t st) {
  RenderSlot& slot = *s->ws->cur;
  if (s->n_edges == 0) return PSDR_OK;
  const CameraK cam = make_camera(camd);
  const int R = static_cast<int>(cfg.primary_edge_hypercube_resolution);
  const int n = edge_lanes(cfg.spp_primary_edge, cam, "primary boundary");
  if (n < 0) return PSDR_ERR_ARG;
  if (slot.b_state.ensure(8ull * std::max(R, 1)) || slot.hc_mass.ensure(4ull * R) || slot.hc_pmf.ensure(4ull * R) ||
      slot.hc_cmf.ensure(4ull * R) || slot.b_tmp.ensure(4ull * (2 * (R / 256 + 2) + 16)))
    return grow_failed();
  hipLaunchKernelGGL(k_seed_states, dim3(blocks(R, BT)), dim3(BT), 0, st, slot.b_state.as<uint64_t>(), R);
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks(R, BT)), dim3(BT), 0, st, slot.hc_mass.as<float>(), R, 0.f);
  if (s->plain) {
    for (uint32_t r = 0; r < cfg.primary_edge_preprocess_rounds; ++r) {
      const int rc = primary_pass(s, cfg, cam, true, R, R, nullptr, nullptr, slot.hc_mass.as<float>(), st);
      if (rc) return rc;
    }
  } else {
    const int group_blocks = blocks(static_cast<int64_t>(R) * PRE_GROUP, PRE_THREADS);
    const size_t pre_threads = static_cast<size_t>(group_blocks) * PRE_THREADS;  // >= the one-per-cell grid
    if (slot.p_spill.ensure(std::max<size_t>(slot.p_spill.bytes, 4ull * (STACK_TOTAL - PRE_STACK) * pre_threads)))
      return grow_failed();
    // one light sample per vertex: a lane per ray of a depth (k_prim_pre_group)
    const bool group = cfg.mis_light_samples == 1;
    for (uint32_t r = 0; r < cfg.primary_edge_preprocess_rounds; ++r) {
      if (group)
        hipLaunchKernelGGL(k_prim_pre_group, dim3(group_blocks), dim3(PRE_THREADS), 0, st, s->view(), cam, R,
                           static_cast<int>(cfg.max_bounce), slot.b_state.as<uint64_t>(), slot.hc_mass.as<float>(),
                           slot.p_spill.as<int>());
      else
        hipLaunchKernelGGL(k_prim_pre, dim3(blocks(R, PRE_THREADS)), dim3(PRE_THREADS), 0, st, s->view(), cam, R,
                           static_cast<int>(cfg.max_bounce), cfg.mis_light_samples, slot.b_state.as<uint64_t>(),
                           slot.hc_mass.as<float>(), slot.p_spill.as<int>());
      PSDR_HIP_CHECK(hipGetLastError());
    }
  }
  int rc = sampler_tables(slot.hc_mass.as<float>(), R, slot.hc_pmf.as<float>(), slot.hc_cmf.as<float>(), nullptr,
                          slot.b_tmp.as<float>(), st);
  if (rc) return rc;
  return primary_pass(s, cfg, cam, false, n, R, slot.hc_pmf.as<float>(), slot.hc_cmf.as<float>(), nullptr, st);
}

int boundary_primary_finish(psdr_scene* s, const psdr_config& cfg, const psdr_camera& camd, const float* dL_dI,
                            float* grad_v, float* cam_acc, hipStream_t st) {
  RenderSlot& slot = *s->ws->cur;
  if (s->n_edges == 0) return PSDR_OK;
  const CameraK cam = make_camera(camd);
  const int n = edge_lanes(cfg.spp_primary_edge, cam, "primary boundary");
  if (n < 0) return PSDR_ERR_ARG;
  if (n == 0) return PSDR_OK;
  const PrimRecs rec = carve_prim(slot.b_rec, n);
  hipLaunchKernelGGL(k_prim_finish<false>, dim3(blocks(n, BT)), dim3(BT), 0, st, s->view(), cam, n, rec,
                     slot.b_LA.as<float>(), slot.b_LB.as<float>(), slot.b_rngB.as<uint64_t>(), nullptr, nullptr, dL_dI,
                     cfg.spp_primary_edge, grad_v, cam_acc);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

int boundary_direct_prepare(psdr_scene* s, const psdr_config& cfg, const psdr_camera& camd, hipStream_t st) {
  RenderSlot& slot = *s->ws->cur;
  if (s->n_edges == 0) return PSDR_OK;
  const CameraK cam = make_camera(camd);
  const SceneView sv = s->view();
  const int r0 = static_cast<int>(cfg.secondary_edge_hypercube_resolution[0]);
  const int r1 = static_cast<int>(cfg.secondary_edge_hypercube_resolution[1]);
  const int r2 = static_cast<int>(cfg.secondary_edge_hypercube_resolution[2]);
  const int64_t cells64 = static_cast<int64_t>(r0) * r1 * r2;
  if (cells64 > 0x7fffffff / 8) {
    set_error(PSDR_ERR_ARG, "direct boundary: too many cells");
    return PSDR_ERR_ARG;
  }
  const int n = edge_lanes(cfg.spp_secondary_edge, cam, "direct boundary");
  if (n < 0) return PSDR_ERR_ARG;
  const int cells = static_cast<int>(cells64);
  const int grid = std::max(1, std::min(blocks(std::max(cells, n), BT), cus() * tuning().direct_bpc.load()));
  if (slot.d_state.ensure(8ull * cells) || slot.d_mass.ensure(4ull * cells) || slot.d_pmf.ensure(4ull * cells) ||
      slot.d_cmf.ensure(4ull * cells) || slot.d_tmp.ensure(4ull * (2 * (cells / 256 + 2) + 16)) ||
      slot.d_spill.ensure(static_cast<size_t>(grid) * BT * (STACK_TOTAL - BSTACK) * 4) ||
      slot.d_keys.ensure(8ull * n) || slot.d_vals.ensure(8ull * n) || slot.d_hist.ensure(4 * radix_hist_words(n)) ||
      slot.d_rec.ensure(48ull * std::max(n, 1)) || slot.d_qidx.ensure(4ull * std::max(cells, n)) ||
      slot.d_qval.ensure(16ull * std::max(cells, n)) || slot.d_qcnt.ensure(4))
    return grow_failed();
  uint64_t* state = slot.d_state.as<uint64_t>();
  int* qidx = slot.d_qidx.as<int>();
  float4* qval = slot.d_qval.as<float4>();
  int* qcnt = slot.d_qcnt.as<int>();
  const int compact = s->plain ? 0 : 1;
  hipLaunchKernelGGL(k_seed_states, dim3(blocks(cells, BT)), dim3(BT), 0, st, state, cells);
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks(cells, BT)), dim3(BT), 0, st, slot.d_mass.as<float>(), cells, 0.f);
  for (uint32_t r = 0; r < cfg.secondary_edge_preprocess_rounds; ++r) {
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, st, qcnt);
    hipLaunchKernelGGL(k_direct_setup<true>, dim3(blocks(cells, BT)), dim3(BT), 0, st, sv, cells, state, r0, r1, r2,
                       nullptr, nullptr, 0u, nullptr, nullptr, nullptr, qidx, qval, qcnt, compact);
    hipLaunchKernelGGL(k_direct<true>, dim3(grid), dim3(BT), 0, st, sv, cam, qcnt, qidx, qval,
                       slot.d_mass.as<float>(), nullptr, slot.d_spill.as<int>());
    PSDR_HIP_CHECK(hipGetLastError());
  }
  int rc = sampler_tables(slot.d_mass.as<float>(), cells, slot.d_pmf.as<float>(), slot.d_cmf.as<float>(), nullptr,
                          slot.d_tmp.as<float>(), st);
  if (rc) return rc;
  if (n == 0) return PSDR_OK;
  // process the lanes in cell order (edge position major): neighbouring lanes
  // of a wave then trace from neighbouring edge points
  const uint32_t* order = nullptr;
  const uint32_t* sorted_cells = nullptr;  // the sort keys: each position's cell
  if (!s->plain) {
    uint32_t* k0 = slot.d_keys.as<uint32_t>();
    uint32_t* v0 = slot.d_vals.as<uint32_t>();
    hipLaunchKernelGGL(k_direct_keys, dim3(blocks(n, BT)), dim3(BT), 0, st, n, cells, slot.d_pmf.as<float>(),
                       slot.d_cmf.as<float>(), k0, v0);
    PSDR_HIP_CHECK(hipGetLastError());
    const int passes = cells <= (1 << 24) ? 3 : 4;
    rc = radix_sort_pairs(k0, v0, k0 + n, v0 + n, slot.d_hist.as<uint32_t>(), n, passes, st);
    if (rc) return rc;
    order = (passes & 1) ? v0 + n : v0;
    sorted_cells = (passes & 1) ? k0 + n : k0;
  }
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, st, qcnt);
  hipLaunchKernelGGL(k_direct_setup<false>, dim3(blocks(n, BT)), dim3(BT), 0, st, sv, n, state, r0, r1, r2,
                     slot.d_pmf.as<float>(), slot.d_cmf.as<float>(), cfg.spp_secondary_edge, slot.d_rec.as<float>(),
                     order, sorted_cells, qidx, qval, qcnt, compact);
  hipLaunchKernelGGL(k_direct<false>, dim3(grid), dim3(BT), 0, st, sv, cam, qcnt, qidx, qval, nullptr,
                     slot.d_rec.as<float>(), slot.d_spill.as<int>());
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

int boundary_direct_finish(psdr_scene* s, const psdr_config& cfg, const psdr_camera& camd, const float* dL_dI,
                           float* grad_v, hipStream_t st) {
  RenderSlot& slot = *s->ws->cur;
  if (s->n_edges == 0) return PSDR_OK;
  const CameraK cam = make_camera(camd);
  const int n = edge_lanes(cfg.spp_secondary_edge, cam, "direct boundary");
  if (n < 0) return PSDR_ERR_ARG;
  if (n == 0) return PSDR_OK;
  hipLaunchKernelGGL(k_direct_scatter, dim3(blocks(n, BT)), dim3(BT), 0, st, s->view(), n, slot.d_rec.as<float>(), dL_dI,
                     grad_v);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

}  // namespace psdr
