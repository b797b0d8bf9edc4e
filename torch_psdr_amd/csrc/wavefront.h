// Wavefront queues shared by the forward render and the boundary estimators.
#pragma once

#include "scene.h"

namespace psdr {

// Path queue (SoA, capacity `cap`): one entry per live path.
struct QueueView {
  uint32_t* lane;
  uint64_t* rng;  // PCG state at the first draw of the pending depth
  float* T;       // [3][cap] throughput
  float* L;       // [3][cap] radiance so far
  int* tri;       // current vertex hit
  float* u;
  float* v;
  float* O;  // [3][cap] origin of the ray that produced the current vertex
  float* J0;
  int cap;
};
constexpr size_t QUEUE_BYTES_PER = 8 + 4 * 5 + 12 * 3;

// Ray queue (xyz interleaved origins/directions) and the traced hits.
struct RaysView {
  float* o;
  float* d;
  float* tl;  // light-sample rays: distance to the sampled light point
  int* tri;
  float* u;
  float* v;
};

struct CacheView {
  int* tri;  // [slots][N]
  float* u;
  float* v;
  int64_t N;
  // per-depth replay state written by the forward render (keep != 0)
  float* T;      // [D][3][N] throughput at the vertex of depth d
  uint64_t* rs;  // [D][N] PCG state before depth d's draws
  float* L;      // [3][N] path radiance before the area jacobian
  int* info;     // [N] depths with light/BSDF work | 256 light camera vertex | 512 radiance not finite
};

// Cameras rendered by one wavefront: lanes [off[j], off[j+1]) of the
// concatenated lane space are camera j's lanes 0 .. N_j - 1 (each camera's
// sampler starts at lane 0, as a render of that camera alone does).  Kept in
// device memory; a queue entry holds its concatenated lane.
constexpr int MAX_CAMS = 4;
struct LaneJobs {
  int n;
  uint32_t off[MAX_CAMS + 1];
  CameraK cam[MAX_CAMS];
  CacheView cache[MAX_CAMS];
};
__device__ __forceinline__ int lane_job(const LaneJobs* J, uint32_t g) {
  int j = 0;
  for (int k = 1; k < MAX_CAMS; ++k)
    if (k < J->n && g >= J->off[k]) j = k;
  return j;
}

// offset of the packed BSDF-sampled ray counts in a run's counter block (runs
// with D < MIS_COUNT_OFF pack them)
constexpr int MIS_COUNT_OFF = 64;

QueueView carve_queue(DevBuf& b, int cap);
RaysView carve_rays(DevBuf& rb, DevBuf& hb, int64_t cap);
RaysView offset_rays(const RaysView& r, int64_t off);

// trace2light (src/path_integrator.cc:80-186) over a queue of paths whose first
// vertex is the hit of a ray from the camera position (rays `r0`, queue `q0`
// with *count0 entries).  Writes radiance (times J0) per lane into lane_L and,
// if lane_rng_out is set, the final PCG state per lane.  Uses q_s[2], r_s[2] as
// ping-pong scratch (capacities >= n paths and n*Kmax rays).
struct WaveArgs {
  CameraK cam;  // the camera (one camera), else cameras come from `jobs`
  int D;
  uint32_t nL, nB;
  int elv, ad_jac, keep;
  CacheView cache;
  float* lane_L;
  uint64_t* lane_rng_out;
  // 128 ints, zeroed by the caller except counters[0] = *count0: [d] paths
  // entering depth d, [MIS_COUNT_OFF + d] packed BSDF-sampled rays of depth d
  int* counters;
  int plain = 0;  // reference launch shape (see psdr_scene::plain)
  DevBuf* spill = nullptr;  // traversal stack overflow area (grown as needed)
  int prio = 0;                    // raise the launches' wave priority (critical path)
  int njobs = 1;                   // cameras in the lane space
  const LaneJobs* jobs = nullptr;  // device table when njobs > 1 (cam / cache unused then)
  // when set (one camera, keep = 0): queue lane fields are positions of the
  // caller's processing order, lane_of[position] is the lane whose PCG stream
  // the path draws from, and lane_L / lane_rng_out are indexed by position
  const uint32_t* lane_of = nullptr;
};
int run_depths(psdr_scene* s, const WaveArgs& a, const QueueView& q0, const RaysView& r0, const QueueView* q_s,
               const RaysView* r_s, int n, hipStream_t st);
int zero_ints(int* p, int n, hipStream_t st);

}  // namespace psdr
