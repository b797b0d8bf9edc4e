// Wavefront forward path tracer (PathIntegrator::render_interior +
// trace2light, src/path_integrator.cc:80-186, 355-372) on gfx950.
//
// Per camera the stages are
//   k_cam      : TEA-seeded PCG32 lane streams, 2 draws, camera rays (all N lanes)
//   k_trace    : closest hit of the ray queue (LDS-stack LBVH traversal)
//   k_shade(d) : resolve depth d-1 (NEE + BSDF-MIS contributions, throughput
//                update from the extension hit) and generate depth d's rays
//                (light-sample, BSDF-sample and extension rays); paths that stay
//                alive are compacted into the next queue with one atomic per workgroup
//   k_film     : per-pixel sum over the pixel's spp lanes in lane order, / spp
// The last depth draws its extension sample (the stream advances exactly as in
// the reference) but traces no extension ray: its hit cannot change L.
// Samples are re-derived from the stored per-depth PCG state in the resolve
// step, so the queue carries only (lane, rng, T, L, hit, origin).
// With keep != 0 every traced hit is also stored per lane (hit cache) for the
// path-replay adjoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "path.h"
#include "render.h"
#include "shading.h"
#include "traverse.h"
#include "wavefront.h"

namespace psdr {

namespace {

constexpr int TRACE_THREADS = TRACE_BLOCK_THREADS;
// render counters of a tile's run (WaveArgs::counters)
constexpr int COUNTERS = 128;

__global__ void k_lane_draws(uint32_t lane0, int n_lanes, int n_draws, uint32_t* out) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lanes) return;
  Rng r = lane_rng(lane0 + static_cast<uint32_t>(l));
  for (int k = 0; k < n_draws; ++k) out[static_cast<int64_t>(l) * n_draws + k] = r.next_u32();
}

__global__ void k_camera_rays(CameraK cam, uint32_t spp, int64_t N, float* o, float* d, int* pix) {
  const int64_t lane = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (lane >= N) return;
  Rng r = lane_rng(static_cast<uint32_t>(lane));
  const float u0 = r.next(), u1 = r.next();
  f3 oo, dd;
  int p;
  camera_ray(cam, static_cast<uint32_t>(lane), spp, u0, u1, &oo, &dd, &p);
  o[3 * lane] = oo.x; o[3 * lane + 1] = oo.y; o[3 * lane + 2] = oo.z;
  d[3 * lane] = dd.x; d[3 * lane + 1] = dd.y; d[3 * lane + 2] = dd.z;
  pix[lane] = p;
}

// Closest hit of rays [0, *count) (or n if count is null); rays SoA-of-xyz
// AoS (ro[3*i+k]); outputs per ray.
// STACK_LDS: traversal stack entries per lane kept in LDS (the rest spill)
// STATS (profiling launches only): per-launch counters of rays, nodes fetched,
// triangles tested (division-free prefilter, then the exact test), closest-hit
// rays whose hit is a near tie or grazes an edge (traverse.h near_tie /
// grazing), and hits; the results equal those of the plain instance.
// Eight waves per SIMD (the gfx950 maximum; the compiler keeps the kernel
// within the SGPR budget by spilling a few scalars to VGPR lanes outside the
// node loop): the traversal waits on dependent node fetches, and the extra wave
// hides more of that latency than the spills cost.
constexpr int TRACE_COUNTERS = 8;
// A timed launch's workgroups stamp their start and end into SPAN_WAYS
// (min, max) pairs picked by workgroup index, reduced on the host: atomics on
// one address are serialised by the memory system (about 11 ns each), and one
// pair for thousands of workgroups would lengthen the launch it measures.
constexpr int SPAN_WAYS = 64;
template <int STACK_LDS, bool STATS>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void k_trace(
    SceneView sv, const float* __restrict__ ro,
                                                         const float* __restrict__ rd, int n,
                                                         const int* __restrict__ count, int K, int seg,
                                                         const float* __restrict__ rtl, int nvis,
                                                         int* __restrict__ otri,
                                                         float* __restrict__ ot, float* __restrict__ ou,
                                                         float* __restrict__ ov, int* __restrict__ spill,
                                                         unsigned long long* __restrict__ timed_bytes,
                                                         unsigned long long* __restrict__ span,
                                                         unsigned long long* __restrict__ counters, int prio,
                                                         int nseek, const int* __restrict__ seek_count) {
  // launches on the step's critical path issue ahead of concurrent kernels'
  // waves on the same SIMD (wave priority, not stream priority)
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ int stack[STACK_LDS * TRACE_THREADS];
  __shared__ unsigned long long cnt_lds[TRACE_COUNTERS];
  if (STATS) {
    if (threadIdx.x < TRACE_COUNTERS) cnt_lds[threadIdx.x] = 0;
    __syncthreads();
  }
  // timed launch: min over span[2w] = earliest workgroup start, max over
  // span[2w + 1] = latest workgroup end (w < SPAN_WAYS) on the device's
  // constant-rate clock, so the difference is
  // the launch's execution time without the time it waited for CUs
  unsigned long long* way = span ? span + 2 * (blockIdx.x % SPAN_WAYS) : nullptr;
  if (way && threadIdx.x == 0) atomicMin(&way[0], static_cast<unsigned long long>(wall_clock64()));
  // rays are K segments of `per` rays each, segment k starting at k * seg
  // (one segment per ray kind, so a wave traces rays of one kind); with
  // seek_count, the one light-seeking segment holds only its first *seek_count
  // entries (the BSDF-sampled rays that can hit an emitting triangle, packed)
  const int per = count ? *count : n;
  const int cnt = per * K;
  const int mseek = seek_count ? *seek_count : per;
  // algorithmic bytes of this launch: ray read (24 B) + tri, u, v (+ t) written
  if (timed_bytes && blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd(timed_bytes, static_cast<unsigned long long>(cnt - (seek_count ? per - mseek : 0)) *
                               (ot ? 40ull : 36ull));
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int ii = gtid; ii < cnt; ii += stride) {
    int i = ii, k = 0;
    if (K > 1) {
      k = ii / per;
      i = k * seg + (ii - k * per);
    }
    if (seek_count && k == nvis && ii - k * per >= mseek) continue;
    const f3 o = f3{ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
    const f3 d = f3{rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
    // segments [0, nvis) are light-sample rays: visibility queries (see
    // closest_hit); [nvis, nvis + nseek) BSDF-sampled rays whose hit matters
    // only when it is an emitting triangle (see light_seeking_hit)
    float tmax = F_INF, tocc = -F_INF;
    if (k < nvis) {
      const float t = rtl[i];
      tmax = vis_tmax(t);
      tocc = vis_tocc(t);
    }
    const bool seek = sv.n_seek >= 0 && k >= nvis && k < nvis + nseek;
    HitRec h;
    if (STATS) {
      RayStats st;
      bool skip = false;
      const HitRec seed = seek_lights(sv.seek_leaf, sv.n_seek, seek, o, d, &skip, st);
      h = skip ? seed
               : closest_hit<STACK_LDS>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, o, d, stack, spill + gtid, stride,
                                        tmax, tocc, st, &seed);
      // visibility queries stop early (no second distance); a light-seeking ray
      // that hit no emitting triangle was not traced
      const bool closest = k >= nvis && !skip;
      atomicAdd(&cnt_lds[0], 1ull);
      atomicAdd(&cnt_lds[1], static_cast<unsigned long long>(st.nodes));
      atomicAdd(&cnt_lds[2], static_cast<unsigned long long>(st.leaves));
      atomicAdd(&cnt_lds[3], static_cast<unsigned long long>(st.exacts));
      if (closest && near_tie(h, st.t2)) atomicAdd(&cnt_lds[4], 1ull);
      if (closest && grazing(h)) atomicAdd(&cnt_lds[5], 1ull);
      if (h.tri >= 0) atomicAdd(&cnt_lds[6], 1ull);
      if (closest) atomicAdd(&cnt_lds[7], 1ull);
    } else {
      NoStats st;
      bool skip = false;
      const HitRec seed = seek_lights(sv.seek_leaf, sv.n_seek, seek, o, d, &skip, st);
      h = skip ? seed
               : closest_hit<STACK_LDS>(sv.node, sv.leaf, sv.n_bvh, sv.n_tris, o, d, stack, spill + gtid, stride,
                                        tmax, tocc, st, &seed);
    }
    otri[i] = h.tri;
    if (ot) ot[i] = h.tri >= 0 ? h.t : -1.f;
    ou[i] = h.u;
    ov[i] = h.v;
  }
  if (span) {
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&way[1], static_cast<unsigned long long>(wall_clock64()));
  }
  if (STATS) {
    __syncthreads();
    if (threadIdx.x < TRACE_COUNTERS) atomicAdd(&counters[threadIdx.x], cnt_lds[threadIdx.x]);
  }
}

// Span slots of timed launches: (start, end) = (max, 0) before a launch.
__global__ void k_span_init(unsigned long long* span, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    span[2 * i] = ~0ull;
    span[2 * i + 1] = 0ull;
  }
}

struct ShadeParams {
  int prio;  // raise the waves' issue priority (the boundary chain on the step's critical path)
  SceneView sv;
  CameraK cam;  // one camera; with njobs > 1 the cameras of `jobs`
  int njobs;
  const LaneJobs* jobs;
  int depth;  // vertex depth being generated (resolve of depth-1 first when > 0)
  int D;
  uint32_t nL, nB;
  int elv;
  int ad_jac;  // renderD semantics: jacobi = A / detach(A)
  int K_in, K_out, Kmax;
  int keep;
  int plain;  // reference launch shape: no compaction, queue entry j stays at slot j
  // the BSDF-sampled rays (nB == 1) that can hit an emitting triangle are packed
  // at the front of their segment; the segment's tl entry of queue slot j holds
  // the index of slot j's ray there, or -1 (no emitting triangle in its way)
  int pack_mis;
  // queue lane fields are positions and lane_of[position] the PCG stream's
  // lane (one camera, no forward state kept); null: lane fields are lanes
  const uint32_t* lane_of;
};

// plain launch shape: a finished path keeps its queue slot with this bit set
constexpr uint32_t LANE_DEAD = 0x80000000u;

__device__ __forceinline__ void store3(float* a, int cap, int j, f3 v) {
  a[j] = v.x;
  a[cap + j] = v.y;
  a[2 * cap + j] = v.z;
}
__device__ __forceinline__ f3 load3(const float* a, int cap, int j) { return f3{a[j], a[cap + j], a[2 * cap + j]}; }
__device__ __forceinline__ void put_ray(RaysView r, int i, f3 o, f3 d) {
  r.o[3 * i] = o.x; r.o[3 * i + 1] = o.y; r.o[3 * i + 2] = o.z;
  r.d[3 * i] = d.x; r.d[3 * i + 1] = d.y; r.d[3 * i + 2] = d.z;
}

__global__ void k_zero_i32(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Camera rays of the n lanes [lane0, lane0 + n) of the (concatenated) lane
// space into queue / ray slots 0 .. n-1.
__global__ void k_cam(CameraK cam, int njobs, const LaneJobs* __restrict__ jobs, uint32_t spp, uint32_t lane0, int n,
                      QueueView q, RaysView rays, int* count_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *count_out = n;
  if (i >= n) return;
  const uint32_t g = lane0 + static_cast<uint32_t>(i);
  uint32_t lane = g;
  const CameraK* c = nullptr;
  if (njobs > 1) {
    const int jb = lane_job(jobs, g);
    lane = g - jobs->off[jb];
    c = &jobs->cam[jb];
  }
  Rng r = lane_rng(lane);
  const float u0 = r.next(), u1 = r.next();
  f3 o, d;
  int p;
  if (njobs > 1) camera_ray(*c, lane, spp, u0, u1, &o, &d, &p);
  else camera_ray(cam, lane, spp, u0, u1, &o, &d, &p);
  put_ray(rays, i, o, d);
  q.lane[i] = g;
  q.rng[i] = r.state;
}

// MULTI: the lanes belong to several cameras (camera and forward-state
// buffers per lane from P.jobs); otherwise to P.cam and cache0.
// LDS: the scene's small tables are staged in LDS (lds_tables_fit).
template <bool MULTI, bool LDS>
__global__ __launch_bounds__(256) void k_shade(ShadeParams P, QueueView qi, const int* __restrict__ count_in,
                                               RaysView rin, QueueView qo, int* __restrict__ count_out,
                                               RaysView rout, CacheView cache0, float* __restrict__ lane_L,
                                               uint64_t* __restrict__ lane_rng_out, int* __restrict__ mis_count_out) {
  if (P.prio) __builtin_amdgcn_s_setprio(3);
  __shared__ int alloc_lds[8];
  SceneView sv = P.sv;
  if constexpr (LDS) {
    __shared__ LdsTables tabs;
    stage_tables(tabs, &sv);
  }
  const int cnt = *count_in;
  const int d = P.depth;
  for (int base = blockIdx.x * blockDim.x; base < cnt; base += gridDim.x * blockDim.x) {
    const int j = base + threadIdx.x;
    bool have = j < cnt;
    bool alive = false;
    bool light0 = false;   // camera vertex on a light, counted with enable_light_visable
    bool dead_in = false;  // plain shape: slot of a path that finished earlier
    uint32_t lane = 0;  // in the concatenated lane space
    uint32_t loc = 0;   // the lane within its camera's render
    Rng rng{0, 0};
    f3 T = splat(1.f), L = splat(0.f);
    float J0 = 1.f;
    Its its;
    float cu = 0.f, cv = 0.f;
    f3 O = P.cam.pos;
    int cam_ortho = P.cam.ortho;
    CacheView cache = cache0;
    if (have && P.plain && (qi.lane[j] & LANE_DEAD)) {
      dead_in = true;
      have = false;
    }
    if (have) {
      lane = qi.lane[j];
      loc = lane;
      if (MULTI) {
        const int jb = lane_job(P.jobs, lane);
        loc = lane - P.jobs->off[jb];
        cache = P.jobs->cache[jb];
        O = P.jobs->cam[jb].pos;
        cam_ortho = P.jobs->cam[jb].ortho;
      }
      rng.state = qi.rng[j];
      rng.inc = lane_inc(P.lane_of ? P.lane_of[lane] : loc);
      if (d == 0) {
        const int ht = rin.tri[j];
        const float hu = rin.u[j], hv = rin.v[j];
        if (P.keep) {
          cache.tri[loc] = ht;
          cache.u[loc] = hu;
          cache.v[loc] = hv;
        }
        // orthographic rays start on the camera plane, one origin per ray
        if (cam_ortho) O = f3{rin.o[3 * j], rin.o[3 * j + 1], rin.o[3 * j + 2]};
        its = reconstruct(sv, ht, hu, hv, O, true);
        cu = hu;
        cv = hv;
        alive = its.valid;
        if (alive && P.ad_jac) J0 = its.area / its.area;
      } else {
        T = load3(qi.T, qi.cap, j);
        L = load3(qi.L, qi.cap, j);
        J0 = qi.J0[j];
        O = load3(qi.O, qi.cap, j);
        const Its ip = reconstruct(sv, qi.tri[j], qi.u[j], qi.v[j], O, true);
        const int rs = qi.cap;  // ray segment stride: ray k of queue entry j is at k * rs + j
        const int64_t cslot = 1 + static_cast<int64_t>(d - 1) * P.Kmax;
        int k = 0;
        // next-event estimation (src/path_integrator.cc:103-131)
        for (uint32_t i = 0; i < P.nL; ++i, ++k) {
          const float ru = rng.next();
          const float rv = rng.next();
          const int ht = rin.tri[k * rs + j];
          const float hu = rin.u[k * rs + j], hv = rin.v[k * rs + j];
          if (P.keep) {
            const int64_t ci = (cslot + k) * cache.N + loc;
            cache.tri[ci] = ht;
            cache.u[ci] = hu;
            cache.v[ci] = hv;
          }
          light_sample_add(sv, ip, ru, rv, ht, hu, hv, P.nL, P.nB, P.ad_jac, T, &L);
        }
        // BSDF sampling with MIS (src/path_integrator.cc:134-160)
        for (uint32_t i = 0; i < P.nB; ++i, ++k) {
          const float ru = rng.next();
          const float rv = rng.next();
          const int mi = P.pack_mis ? __float_as_int(rin.tl[k * rs + j]) : j;
          const int ht = mi >= 0 ? rin.tri[k * rs + mi] : -1;
          const float hu = mi >= 0 ? rin.u[k * rs + mi] : -1.f, hv = mi >= 0 ? rin.v[k * rs + mi] : -1.f;
          if (P.keep) {
            const int64_t ci = (cslot + k) * cache.N + loc;
            cache.tri[ci] = ht;
            cache.u[ci] = hu;
            cache.v[ci] = hv;
          }
          bsdf_sample_add(sv, ip, ru, rv, ht, hu, hv, P.nL, P.nB, P.ad_jac, T, &L);
        }
        // extension (src/path_integrator.cc:163-181)
        {
          const float ru = rng.next();
          const float rv = rng.next();
          if (d - 1 < P.D - 1) {
            const int ht = rin.tri[k * rs + j];
            const float hu = rin.u[k * rs + j], hv = rin.v[k * rs + j];
            if (P.keep) {
              const int64_t ci = (cslot + k) * cache.N + loc;
              cache.tri[ci] = ht;
              cache.u[ci] = hu;
              cache.v[ci] = hv;
            }
            alive = extend_path(sv, ip, ru, rv, ht, hu, hv, P.ad_jac, &its, &T, &O);
            cu = hu;
            cv = hv;
          }
        }
      }
      // generate depth d at `its` (src/path_integrator.cc:92-99 first)
      if (alive && d < P.D) {
        if (its.is_light) {
          if (P.elv && d == 0) {
            const int l = its.light;
            L = L + T * light_radiance(sv, l, its);
            light0 = true;
          }
          alive = false;
        }
      } else {
        alive = false;
      }
    }
    int slot = j;
    if (P.plain) {
      if (j == 0) *count_out = cnt;
      if (dead_in || (have && !alive)) {
        // keep the slot, marked finished, with rays that miss at once
        qo.lane[j] = (have ? lane : qi.lane[j]) | LANE_DEAD;
        const f3 nan3 = splat(__builtin_nanf(""));
        for (int k = 0; k < P.K_out; ++k) put_ray(rout, k * qo.cap + j, nan3, nan3);
      }
    } else {
      slot = block_alloc(alive, count_out, alloc_lds);
    }
    bool mis_want = false;  // packed BSDF-sampled ray to write
    f3 mis_d = splat(0.f);
    uint64_t rng_start = 0;
    if (have && alive) {
      rng_start = rng.state;
      if (P.keep) {
        const int64_t N = cache.N;
        cache.T[(3ll * d) * N + loc] = T.x;
        cache.T[(3ll * d + 1) * N + loc] = T.y;
        cache.T[(3ll * d + 2) * N + loc] = T.z;
        cache.rs[d * N + loc] = rng_start;
      }
      const int rs = qo.cap;
      int k = 0;
      for (uint32_t i = 0; i < P.nL; ++i, ++k) {
        const float ru = rng.next();
        const float rv = rng.next();
        f3 dir;
        float tl;
        light_sample_ray(sv, its.p, ru, rv, &dir, &tl);
        put_ray(rout, k * rs + slot, its.p, dir);
        rout.tl[k * rs + slot] = tl;
      }
      for (uint32_t i = 0; i < P.nB; ++i, ++k) {
        const float ru = rng.next();
        const float rv = rng.next();
        const Sampled smp = mat_sample(sv, its, ru, rv);
        if (P.pack_mis) {
          bool skip;
          NoStats nst;
          (void)seek_lights(sv.seek_leaf, sv.n_seek, true, its.p, smp.wo, &skip, nst);
          mis_want = !skip;
          mis_d = smp.wo;
        } else {
          put_ray(rout, k * rs + slot, its.p, smp.wo);
        }
      }
    }
    if (P.pack_mis) {
      const int seg = static_cast<int>(P.nL) * qo.cap;
      const int m = block_alloc(mis_want, mis_count_out, alloc_lds);
      if (mis_want) put_ray(rout, seg + m, its.p, mis_d);
      if (have && alive) rout.tl[seg + slot] = __int_as_float(mis_want ? m : -1);
    }
    if (have && alive) {
      const int rs = qo.cap;
      const int k = static_cast<int>(P.nL + P.nB);
      {
        const float ru = rng.next();
        const float rv = rng.next();
        if (d < P.D - 1) {
          const Sampled smp = mat_sample(sv, its, ru, rv);
          put_ray(rout, k * rs + slot, its.p, smp.wo);
        }
      }
      qo.lane[slot] = lane;
      qo.rng[slot] = rng_start;
      store3(qo.T, qo.cap, slot, T);
      store3(qo.L, qo.cap, slot, L);
      store3(qo.O, qo.cap, slot, O);
      qo.J0[slot] = J0;
      qo.tri[slot] = its.tri;
      qo.u[slot] = cu;
      qo.v[slot] = cv;
    } else if (have) {
      const bool finite = isfinite(sqnorm(L));
      if (P.keep) {
        const int64_t N = cache.N;
        cache.L[loc] = L.x;
        cache.L[N + loc] = L.y;
        cache.L[2 * N + loc] = L.z;
        cache.info[loc] = d | (light0 ? 256 : 0) | (finite ? 0 : 512);
      }
      if (!finite) L = splat(0.f);
      L = L * J0;
      lane_L[3 * static_cast<int64_t>(lane)] = L.x;
      lane_L[3 * static_cast<int64_t>(lane) + 1] = L.y;
      lane_L[3 * static_cast<int64_t>(lane) + 2] = L.z;
      if (lane_rng_out) lane_rng_out[lane] = rng.state;
    }
  }
}

// Film::add + n_pass averaging (src/camera.h:75-87, src/path_integrator.cc:385-400).
// A block stages a tile of whole pixels' lane radiance through LDS (coalesced
// loads), then each (pixel, channel) thread sums its spp lanes in lane order.
constexpr int FILM_LANES = 1024;
__global__ __launch_bounds__(256) void k_film(const float* __restrict__ lane_L, int P_, uint32_t spp,
                                              uint32_t n_pass, float* __restrict__ out) {
  __shared__ float tile[3 * FILM_LANES];
  const int ppb = spp <= FILM_LANES ? FILM_LANES / static_cast<int>(spp) : 1;  // pixels per block
  const int p0 = blockIdx.x * ppb;
  const int np = min(ppb, P_ - p0);
  if (np <= 0) return;
  const int tid = threadIdx.x;
  const int pc = tid / 3, c = tid % 3;  // this thread's (pixel, channel) within the tile
  float s = 0.f;
  for (int64_t l0 = 0; l0 < static_cast<int64_t>(np) * spp; l0 += FILM_LANES) {
    const int64_t base = static_cast<int64_t>(p0) * spp + l0;
    const int nl = static_cast<int>(min(static_cast<int64_t>(FILM_LANES), static_cast<int64_t>(np) * spp - l0));
    __syncthreads();
    for (int i = tid; i < 3 * nl; i += blockDim.x) tile[i] = lane_L[3 * base + i];
    __syncthreads();
    if (spp <= FILM_LANES) {
      for (int q = pc; q < np; q += blockDim.x / 3) {
        if (tid >= 3 * (blockDim.x / 3)) break;
        float acc = 0.f;
        for (uint32_t k = 0; k < spp; ++k) acc = acc + tile[3 * (q * static_cast<int>(spp) + static_cast<int>(k)) + c];
        if (spp > 1) acc = acc / static_cast<float>(spp);
        float fin = 0.f;
        for (uint32_t n = 0; n < n_pass; ++n) fin = fin + acc;
        if (n_pass > 1) fin = fin / static_cast<float>(n_pass);
        out[3 * (p0 + q) + c] = fin;
      }
    } else if (tid < 3) {
      for (int k = 0; k < nl; ++k) s = s + tile[3 * k + c];
    }
  }
  if (spp > FILM_LANES && tid < 3) {
    float acc = s / static_cast<float>(spp);
    float fin = 0.f;
    for (uint32_t n = 0; n < n_pass; ++n) fin = fin + acc;
    if (n_pass > 1) fin = fin / static_cast<float>(n_pass);
    out[3 * p0 + c] = fin;
  }
}

inline int blocks(int64_t n, int t) { return static_cast<int>((n + t - 1) / t); }

int g_cus = 0;

}  // namespace

int num_cus() {
  if (g_cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_cus = prop.multiProcessorCount;
    if (g_cus <= 0) g_cus = 256;
  }
  return g_cus;
}

Tuning& tuning() {
  static Tuning t;
  return t;
}

bool trace_stack_ok(int entries) { return entries == 8 || entries == 16 || entries == 32; }

QueueView carve_queue(DevBuf& b, int cap) {
  QueueView q;
  char* p = static_cast<char*>(b.p);
  q.cap = cap;
  q.rng = reinterpret_cast<uint64_t*>(p);
  p += sizeof(uint64_t) * cap;
  q.lane = reinterpret_cast<uint32_t*>(p);
  p += 4 * static_cast<size_t>(cap);
  q.tri = reinterpret_cast<int*>(p);
  p += 4 * static_cast<size_t>(cap);
  q.u = reinterpret_cast<float*>(p);
  p += 4 * static_cast<size_t>(cap);
  q.v = reinterpret_cast<float*>(p);
  p += 4 * static_cast<size_t>(cap);
  q.J0 = reinterpret_cast<float*>(p);
  p += 4 * static_cast<size_t>(cap);
  q.T = reinterpret_cast<float*>(p);
  p += 12 * static_cast<size_t>(cap);
  q.L = reinterpret_cast<float*>(p);
  p += 12 * static_cast<size_t>(cap);
  q.O = reinterpret_cast<float*>(p);
  return q;
}

RaysView carve_rays(DevBuf& rb, DevBuf& hb, int64_t cap) {
  RaysView r;
  r.o = rb.as<float>();
  r.d = r.o + 3 * cap;
  r.tl = r.d + 3 * cap;
  r.tri = hb.as<int>();
  r.u = reinterpret_cast<float*>(r.tri + cap);
  r.v = r.u + cap;
  return r;
}

RaysView offset_rays(const RaysView& r, int64_t off) {
  RaysView o;
  o.o = r.o + 3 * off;
  o.d = r.d + 3 * off;
  o.tl = r.tl + off;
  o.tri = r.tri + off;
  o.u = r.u + off;
  o.v = r.v + off;
  return o;
}

int zero_ints(int* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_i32, dim3(blocks(n, 256)), dim3(256), 0, st, p, n);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

namespace {
// Optional live timing of every k_trace launch (HIP events on the launch
// stream) plus the launches' algorithmic bytes counted on the device.
struct TraceTimer {
  std::mutex mu;
  bool on = false;
  std::vector<hipEvent_t> ev;  // (start, end) pairs
  size_t used = 0;             // pairs recorded since the last start
  int64_t dropped = 0;         // launches beyond the event pool
  unsigned long long* bytes = nullptr;
  unsigned long long* span = nullptr;  // [pair][way][start, end] device clock of each launch
  double clock_khz = 0.0;
  bool count_on = false;                   // launches use the counting instance
  unsigned long long* counters = nullptr;  // TRACE_COUNTERS totals
};
TraceTimer& trace_timer() {
  static TraceTimer t;
  return t;
}
constexpr size_t TIMER_PAIRS = 4096;
}  // namespace

int trace_rays(const SceneView& sv, const float* ro, const float* rd, int n, const int* count_dev, int K, int seg,
               const float* tl, int nvis, int* tri, float* t, float* u, float* v, int* spill, int grid,
               hipStream_t st, int stack_lds, int prio, int nseek, const int* seek_count) {
  TraceTimer& tt = trace_timer();
  std::lock_guard<std::mutex> lock(tt.mu);
  const bool timed = tt.on && tt.used < TIMER_PAIRS;
  if (tt.on && !timed) ++tt.dropped;
  if (timed) PSDR_HIP_CHECK(hipEventRecord(tt.ev[2 * tt.used], st));
  unsigned long long* tb = timed ? tt.bytes : nullptr;
  unsigned long long* sp = timed ? tt.span + 2 * SPAN_WAYS * tt.used : nullptr;
  unsigned long long* cn = tt.count_on ? tt.counters : nullptr;
#define PSDR_LAUNCH_TRACE(S, STATS)                                                                                  \
  hipLaunchKernelGGL((k_trace<S, STATS>), dim3(grid), dim3(TRACE_THREADS), 0, st, sv, ro, rd, n, count_dev, K, seg, tl, \
                     nvis, tri, t, u, v, spill, tb, sp, cn, prio, nseek, seek_count)
  switch (stack_lds ? stack_lds : tuning().trace_stack.load()) {
    case 8:
      if (cn) PSDR_LAUNCH_TRACE(8, true);
      else PSDR_LAUNCH_TRACE(8, false);
      break;
    case 32:
      if (cn) PSDR_LAUNCH_TRACE(32, true);
      else PSDR_LAUNCH_TRACE(32, false);
      break;
    default:
      if (cn) PSDR_LAUNCH_TRACE(16, true);
      else PSDR_LAUNCH_TRACE(16, false);
      break;
  }
#undef PSDR_LAUNCH_TRACE
  PSDR_HIP_CHECK(hipGetLastError());
  if (timed) {
    PSDR_HIP_CHECK(hipEventRecord(tt.ev[2 * tt.used + 1], st));
    ++tt.used;
  }
  return PSDR_OK;
}

int trace_timing(int enable, double* total_ms, int64_t* launches, int64_t* bytes, double* busy_ms) {
  TraceTimer& tt = trace_timer();
  std::lock_guard<std::mutex> lock(tt.mu);
  if (enable) {
    if (tt.ev.empty()) {
      int dev = 0, khz = 0;
      PSDR_HIP_CHECK(hipGetDevice(&dev));
      PSDR_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
      if (khz <= 0) {
        set_error(PSDR_ERR_HIP, "trace_timing: the device reports no wall-clock rate");
        return PSDR_ERR_HIP;
      }
      tt.clock_khz = khz;
      tt.ev.resize(2 * TIMER_PAIRS);
      for (auto& e : tt.ev) PSDR_HIP_CHECK(hipEventCreate(&e));
      PSDR_HIP_CHECK(hipMalloc(&tt.bytes, sizeof(unsigned long long)));
      PSDR_HIP_CHECK(hipMalloc(&tt.span, 2 * SPAN_WAYS * TIMER_PAIRS * sizeof(unsigned long long)));
    }
    PSDR_HIP_CHECK(hipDeviceSynchronize());
    PSDR_HIP_CHECK(hipMemset(tt.bytes, 0, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_span_init, dim3(blocks(TIMER_PAIRS * SPAN_WAYS, 256)), dim3(256), 0, nullptr, tt.span,
                       static_cast<int>(TIMER_PAIRS * SPAN_WAYS));
    PSDR_HIP_CHECK(hipGetLastError());
    PSDR_HIP_CHECK(hipDeviceSynchronize());
    tt.used = 0;
    tt.dropped = 0;
    tt.on = true;
    return PSDR_OK;
  }
  if (!tt.on) {
    set_error(PSDR_ERR_STATE, "trace_timing: timing was not started");
    return PSDR_ERR_STATE;
  }
  tt.on = false;
  PSDR_HIP_CHECK(hipDeviceSynchronize());
  double ms = 0.0;
  for (size_t i = 0; i < tt.used; ++i) {
    float e = 0.f;
    PSDR_HIP_CHECK(hipEventElapsedTime(&e, tt.ev[2 * i], tt.ev[2 * i + 1]));
    ms += e;
  }
  unsigned long long b = 0;
  PSDR_HIP_CHECK(hipMemcpy(&b, tt.bytes, sizeof(b), hipMemcpyDeviceToHost));
  std::vector<unsigned long long> span(2 * SPAN_WAYS * tt.used);
  if (tt.used) PSDR_HIP_CHECK(hipMemcpy(span.data(), tt.span, span.size() * sizeof(unsigned long long),
                                        hipMemcpyDeviceToHost));
  double ticks = 0.0;
  for (size_t i = 0; i < tt.used; ++i) {
    unsigned long long t0 = ~0ull, t1 = 0ull;  // ways no workgroup used keep (max, 0)
    for (int w = 0; w < SPAN_WAYS; ++w) {
      t0 = std::min(t0, span[2 * (SPAN_WAYS * i + w)]);
      t1 = std::max(t1, span[2 * (SPAN_WAYS * i + w) + 1]);
    }
    if (t1 >= t0) ticks += static_cast<double>(t1 - t0);
  }
  if (tt.dropped) {
    set_error(PSDR_ERR_STATE, "trace_timing: more launches than the event pool holds");
    return PSDR_ERR_STATE;
  }
  if (busy_ms) *busy_ms = ticks / tt.clock_khz;
  if (total_ms) *total_ms = ms;
  if (launches) *launches = static_cast<int64_t>(tt.used);
  if (bytes) *bytes = static_cast<int64_t>(b);
  return PSDR_OK;
}

int trace_counters(int enable, int64_t* out, int n_out) {
  TraceTimer& tt = trace_timer();
  std::lock_guard<std::mutex> lock(tt.mu);
  if (enable) {
    if (!tt.counters) PSDR_HIP_CHECK(hipMalloc(&tt.counters, TRACE_COUNTERS * sizeof(unsigned long long)));
    PSDR_HIP_CHECK(hipDeviceSynchronize());
    PSDR_HIP_CHECK(hipMemset(tt.counters, 0, TRACE_COUNTERS * sizeof(unsigned long long)));
    PSDR_HIP_CHECK(hipDeviceSynchronize());
    tt.count_on = true;
    return PSDR_OK;
  }
  if (!tt.count_on) {
    set_error(PSDR_ERR_STATE, "trace_counters: counting was not started");
    return PSDR_ERR_STATE;
  }
  tt.count_on = false;
  PSDR_HIP_CHECK(hipDeviceSynchronize());
  unsigned long long c[TRACE_COUNTERS];
  PSDR_HIP_CHECK(hipMemcpy(c, tt.counters, sizeof(c), hipMemcpyDeviceToHost));
  for (int i = 0; i < n_out && i < TRACE_COUNTERS; ++i) out[i] = static_cast<int64_t>(c[i]);
  return PSDR_OK;
}

int trace_grid(int64_t max_rays) {
  const int g = std::max(1, std::min(blocks(max_rays, TRACE_THREADS), num_cus() * tuning().trace_bpc.load()));
  return g;
}

size_t spill_bytes(int grid) {
  return static_cast<size_t>(grid) * TRACE_THREADS * (STACK_TOTAL - TRACE_STACK_MIN) * 4;
}

int lane_draws(uint32_t lane0, int n_lanes, int n_draws, uint32_t* out, hipStream_t st) {
  if (n_lanes <= 0) return PSDR_OK;
  hipLaunchKernelGGL(k_lane_draws, dim3(blocks(n_lanes, 256)), dim3(256), 0, st, lane0, n_lanes, n_draws, out);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

int camera_rays(const CameraK& cam, uint32_t spp, float* o, float* d, int* pix, hipStream_t st) {
  const int64_t N = static_cast<int64_t>(spp) * cam.H * cam.W;
  if (N <= 0) return PSDR_OK;
  hipLaunchKernelGGL(k_camera_rays, dim3(blocks(N, 256)), dim3(256), 0, st, cam, spp, N, o, d, pix);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

int run_depths(psdr_scene* s, const WaveArgs& a, const QueueView& q0, const RaysView& r0, const QueueView* q_s,
               const RaysView* r_s, int n, hipStream_t st) {
  const SceneView sv = s->view();
  if (a.lane_of && (a.keep || a.njobs > 1)) {
    set_error(PSDR_ERR_ARG, "run_depths: position-indexed lanes need one camera and no kept forward state");
    return PSDR_ERR_ARG;
  }
  const int Kmax = static_cast<int>(a.nL + a.nB + 1);
  const int64_t ray_cap = static_cast<int64_t>(n) * Kmax;
  const int tgrid = trace_grid(ray_cap);
  DevBuf& spill = a.spill ? *a.spill : s->ws->spill;
  if (spill.ensure(std::max(spill.bytes, spill_bytes(tgrid)))) return grow_failed();
  const int sgrid = std::max(1, std::min(blocks(n, 256), num_cus() * tuning().shade_bpc.load()));
  QueueView qin = q0;
  RaysView rin = r0;
  // ray k of queue entry j is at k * cap + j (k_shade); the queues of one run
  // share a capacity >= n
  const int cap = q_s[0].cap;
  // one BSDF-sampled ray per vertex and few emitting triangles: only the rays
  // that can hit one are traced, packed (their counts after the depth counts)
  const int pack = !a.plain && sv.n_seek >= 0 && a.nB == 1 && a.D < MIS_COUNT_OFF ? 1 : 0;
  for (int d = 0; d <= a.D; ++d) {
    const QueueView qout = q_s[d & 1];
    const RaysView rout = r_s[d & 1];
    ShadeParams P;
    P.prio = a.prio;
    P.sv = sv;
    P.cam = a.cam;
    P.njobs = a.njobs;
    P.jobs = a.jobs;
    P.depth = d;
    P.D = a.D;
    P.nL = a.nL;
    P.nB = a.nB;
    P.elv = a.elv;
    P.ad_jac = a.ad_jac;
    P.K_in = d == 0 ? 1 : static_cast<int>(a.nL + a.nB) + ((d - 1 < a.D - 1) ? 1 : 0);
    P.K_out = d < a.D ? static_cast<int>(a.nL + a.nB) + ((d < a.D - 1) ? 1 : 0) : 0;
    P.Kmax = Kmax;
    P.keep = a.keep;
    P.plain = a.plain;
    P.pack_mis = pack;
    P.lane_of = a.lane_of;
    int* mis_count = pack ? a.counters + MIS_COUNT_OFF + d : nullptr;
    const bool lds = lds_tables_fit(sv);
    auto kern = a.njobs > 1 ? (lds ? k_shade<true, true> : k_shade<true, false>)
                            : (lds ? k_shade<false, true> : k_shade<false, false>);
    hipLaunchKernelGGL(kern, dim3(sgrid), dim3(256), 0, st, P, qin, a.counters + d, rin, qout, a.counters + d + 1,
                       rout, a.cache, a.lane_L, a.lane_rng_out, mis_count);
    PSDR_HIP_CHECK(hipGetLastError());
    if (d < a.D && P.K_out > 0) {
      if (a.plain) {
        // one full-width launch per ray kind, each followed by a host sync
        const int g1 = trace_grid(n);
        for (int k = 0; k < P.K_out; ++k) {
          const int64_t o = static_cast<int64_t>(k) * cap;
          const int kind_vis = k < static_cast<int>(a.nL) ? 1 : 0;
          const int kind_seek = !kind_vis && k < static_cast<int>(a.nL + a.nB) ? 1 : 0;
          const int rc = trace_rays(sv, rout.o + 3 * o, rout.d + 3 * o, n, nullptr, 1, 0, rout.tl + o, kind_vis,
                                    rout.tri + o, nullptr, rout.u + o, rout.v + o, spill.as<int>(), g1, st, 0, a.prio,
                                    kind_seek);
          if (rc) return rc;
          PSDR_HIP_CHECK(hipStreamSynchronize(st));
        }
      } else {
        // segments: nL light-sample rays, nB BSDF-sampled rays, the extension
        const int rc = trace_rays(sv, rout.o, rout.d, n, a.counters + d + 1, P.K_out, cap, rout.tl,
                                  static_cast<int>(a.nL), rout.tri, nullptr, rout.u, rout.v, spill.as<int>(), tgrid,
                                  st, 0, a.prio, static_cast<int>(a.nB), mis_count);
        if (rc) return rc;
      }
    }
    qin = qout;
    rin = rout;
  }
  return PSDR_OK;
}

// The camera's forward-state buffers (hit cache and replay state) in `slot`.
static int carve_cache(RenderSlot* slot, int64_t N, int D, int Kmax, CacheView* cache) {
  const int64_t slots = 1 + static_cast<int64_t>(D) * Kmax;
  *cache = CacheView{nullptr, nullptr, nullptr, N, nullptr, nullptr, nullptr, nullptr};
  if (!slot) return PSDR_OK;
  if (slot->hit_cache.ensure(12ull * slots * N) || slot->replay.ensure((20ull * std::max(D, 1) + 16ull) * N))
    return grow_failed();
  cache->tri = slot->hit_cache.as<int>();
  cache->u = reinterpret_cast<float*>(cache->tri + slots * N);
  cache->v = cache->u + slots * N;
  cache->rs = slot->replay.as<uint64_t>();
  cache->T = reinterpret_cast<float*>(cache->rs + static_cast<int64_t>(std::max(D, 1)) * N);
  cache->L = cache->T + 3ll * std::max(D, 1) * N;
  cache->info = reinterpret_cast<int*>(cache->L + 3ll * N);
  return PSDR_OK;
}

static_assert(MAX_CAMS == PSDR_MAX_CAMERAS, "cameras per wavefront");

__global__ void k_set_jobs(LaneJobs J, LaneJobs* dst) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = J;
}

// Cameras in one wavefront: their lanes are concatenated and processed in tiles
// of at most `tile` lanes (tuning; 0 = all in one), so that the queue, ray and
// hit buffers hold one tile while the per-lane radiance and the forward states
// hold every lane.  Each tile uses its own block of counters.
int render_forward(psdr_scene* s, const psdr_config& cfg, const psdr_camera* camd, int n_cams, float* const* outs,
                   int keep, hipStream_t st) {
  if (n_cams < 1 || n_cams > MAX_CAMS) {
    set_error(PSDR_ERR_ARG, "render: 1 to 4 cameras per wavefront");
    return PSDR_ERR_ARG;
  }
  LaneJobs J;
  std::memset(&J, 0, sizeof(J));
  J.n = n_cams;
  int64_t total = 0;
  for (int c = 0; c < n_cams; ++c) {
    J.cam[c] = make_camera(camd[c]);
    const int64_t n64 = static_cast<int64_t>(cfg.spp_interior) * J.cam[c].H * J.cam[c].W;
    if (n64 <= 0 || n64 > 0x7fffffff / 4 || cfg.max_bounce > 60) {
      set_error(PSDR_ERR_ARG, "render: spp*H*W out of range");
      return PSDR_ERR_ARG;
    }
    J.off[c] = static_cast<uint32_t>(total);
    total += n64;
  }
  J.off[n_cams] = static_cast<uint32_t>(total);
  if (total > 0x7fffffff) {
    set_error(PSDR_ERR_ARG, "render: the cameras of one wavefront hold more than 2^31 - 1 lanes");
    return PSDR_ERR_ARG;
  }
  const int D = static_cast<int>(cfg.max_bounce);
  const uint32_t nL = cfg.mis_light_samples, nB = cfg.mis_bsdf_samples;
  const int Kmax = static_cast<int>(nL + nB + 1);
  // a previous boundary prepass reads buffers this render rewrites or the
  // caller may change afterwards: order it before this render
  int rcj = aux_join(s, st);
  if (rcj) return rcj;
  // renderD: each camera's forward state goes to its (config, camera) slot
  RenderSlot* slots[MAX_CAMS] = {};
  if (keep) {
    for (int c = 0; c < n_cams; ++c) {
      slots[c] = s->ws->claim_slot(s, cfg, camd[c]);
      slots[c]->have_cache = false;
      slots[c]->bprep = false;
    }
  }
  const bool prep = keep && !s->plain && s->n_edges > 0 && (cfg.spp_primary_edge > 0 || cfg.spp_secondary_edge > 0);
  if (prep) {
    // renderD: the boundary terms' dL/dI-independent work goes to the
    // auxiliary streams, forked here so that it depends on nothing of this
    // render and overlaps it and the caller's loss.  It is enqueued after this
    // render's own launches (below): kernels start in the order the host
    // submits them, and the interior chain (render, loss, adjoint) is the
    // step's longest, so it must not wait behind the boundary chains' launches.
    const int rca = s->ws->aux.ensure();
    if (rca) return rca;
    PSDR_HIP_CHECK(hipEventRecord(s->ws->aux.fork, st));
  }
  // tiles (the reference launch shape renders every lane at once); a tile's
  // rays are indexed as 3 * i in 32-bit arithmetic, i < tile * Kmax
  int64_t T = tuning().tile_lanes.load();
  if (T <= 0 || T > total || s->plain) T = total;
  T = std::min<int64_t>(T, 0x7fffffff / (3ll * std::max(Kmax, 1)));
  const int n_tiles = static_cast<int>((total + T - 1) / T);
  const int64_t ray_cap = T * std::max(Kmax, 1);
  if (s->ws->lane_L.ensure(12ull * total) || s->ws->q_a.ensure(QUEUE_BYTES_PER * T) ||
      s->ws->q_b.ensure(QUEUE_BYTES_PER * T) || s->ws->rays.ensure(RAY_BYTES_PER * ray_cap) ||
      s->ws->hits.ensure(12ull * ray_cap) || s->ws->rays2.ensure(RAY_BYTES_PER * ray_cap) ||
      s->ws->hits2.ensure(12ull * ray_cap) || s->ws->counters_r.ensure(4ull * COUNTERS * n_tiles) ||
      s->ws->jobs.ensure(sizeof(LaneJobs)))
    return grow_failed();
  const int tgrid = trace_grid(ray_cap);
  if (s->ws->spill.ensure(spill_bytes(tgrid))) return grow_failed();
  for (int c = 0; c < n_cams; ++c) {
    const int rc = carve_cache(slots[c], static_cast<int64_t>(J.off[c + 1]) - J.off[c], D, Kmax, &J.cache[c]);
    if (rc) return rc;
  }
  LaneJobs* jobs_dev = s->ws->jobs.as<LaneJobs>();
  if (n_cams > 1) {
    hipLaunchKernelGGL(k_set_jobs, dim3(1), dim3(1), 0, st, J, jobs_dev);
    PSDR_HIP_CHECK(hipGetLastError());
  }
  const int Ti = static_cast<int>(T);
  QueueView qa = carve_queue(s->ws->q_a, Ti), qb = carve_queue(s->ws->q_b, Ti);
  // the shade kernel reads the hits of the previous depth while writing the
  // next depth's rays: rays and hits are double-buffered
  RaysView rays_a = carve_rays(s->ws->rays, s->ws->hits, ray_cap);
  RaysView rays_b = carve_rays(s->ws->rays2, s->ws->hits2, ray_cap);
  int* counters = s->ws->counters_r.as<int>();
  hipLaunchKernelGGL(k_zero_i32, dim3(blocks(COUNTERS * n_tiles, 256)), dim3(256), 0, st, counters,
                     COUNTERS * n_tiles);
  const SceneView sv = s->view();
  float* lane_L = s->ws->lane_L.as<float>();
  WaveArgs wa;
  wa.cam = J.cam[0];
  wa.D = D;
  wa.nL = nL;
  wa.nB = nB;
  wa.elv = cfg.enable_light_visable ? 1 : 0;
  wa.ad_jac = keep ? 1 : 0;
  wa.keep = keep;
  wa.cache = J.cache[0];
  wa.lane_L = lane_L;
  wa.lane_rng_out = nullptr;
  wa.plain = s->plain ? 1 : 0;
  wa.njobs = n_cams;
  wa.jobs = n_cams > 1 ? jobs_dev : nullptr;
  const QueueView qs[2] = {qb, qa};
  const RaysView rs[2] = {rays_b, rays_a};
  for (int t = 0; t < n_tiles; ++t) {
    const int64_t l0 = static_cast<int64_t>(t) * T;
    const int n = static_cast<int>(std::min<int64_t>(T, total - l0));
    int* cnt = counters + COUNTERS * t;
    // stage 0: camera rays + primary hits
    hipLaunchKernelGGL(k_cam, dim3(blocks(n, 256)), dim3(256), 0, st, J.cam[0], n_cams, wa.jobs, cfg.spp_interior,
                       static_cast<uint32_t>(l0), n, qa, rays_a, cnt);
    PSDR_HIP_CHECK(hipGetLastError());
    int rc = trace_rays(sv, rays_a.o, rays_a.d, n, cnt, 1, Ti, nullptr, 0, rays_a.tri, nullptr, rays_a.u, rays_a.v,
                        s->ws->spill.as<int>(), tgrid, st);
    if (rc) return rc;
    if (s->plain) PSDR_HIP_CHECK(hipStreamSynchronize(st));
    wa.counters = cnt;
    rc = run_depths(s, wa, qa, rays_a, qs, rs, n, st);
    if (rc) return rc;
  }
  const int ppb = cfg.spp_interior <= static_cast<uint32_t>(FILM_LANES) ? FILM_LANES / static_cast<int>(cfg.spp_interior)
                                                                        : 1;
  for (int c = 0; c < n_cams; ++c) {
    const int P_ = J.cam[c].H * J.cam[c].W;
    hipLaunchKernelGGL(k_film, dim3(blocks(P_, ppb)), dim3(256), 0, st, lane_L + 3ll * J.off[c], P_, cfg.spp_interior,
                       cfg.n_pass, outs[c]);
    PSDR_HIP_CHECK(hipGetLastError());
  }
  if (prep) {
    for (int c = 0; c < n_cams; ++c) {
      s->ws->cur = slots[c];
      const int rcp = boundary_prepare_async(s, cfg, camd[c]);
      if (rcp) return rcp;
    }
  }
  s->ws->last_owner = s;
  s->ws->last_N = static_cast<uint64_t>(total);
  s->ws->last_K = Kmax;
  s->ws->last_D = D;
  s->ws->last_tiles = n_tiles;
  s->ws->last_cfg = cfg;
  s->ws->last_cam = camd[n_cams - 1];
  for (int c = 0; c < n_cams; ++c) {
    if (!slots[c]) continue;
    slots[c]->N = J.off[c + 1] - J.off[c];
    slots[c]->K = Kmax;
    slots[c]->D = D;
    slots[c]->cfg = cfg;
    slots[c]->cam = camd[c];
    slots[c]->have_cache = true;
  }
  if (keep) s->ws->cur = slots[n_cams - 1];
  return PSDR_OK;
}

}  // namespace psdr
