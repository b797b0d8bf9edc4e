// Host-side entry points of the renderer kernels (render.hip, adjoint.hip).
#pragma once

#include <atomic>

#include "scene.h"

namespace psdr {
// Launch shapes chosen at run time (psdr_set_tuning), process-wide: grids in
// blocks per CU, and the traversal stack entries kept in LDS per lane.
struct Tuning {
  std::atomic<int> trace_bpc{16}, trace_stack{16}, shade_bpc{8}, adjoint_bpc{4}, direct_bpc{8};
  std::atomic<int> tile_lanes{0};  // lanes per wavefront tile of a forward render, 0 = all
};
Tuning& tuning();
constexpr int TUNE_BPC_MAX = 64;
// LDS stack depths with a k_trace instance; spill areas are sized for the smallest
constexpr int TRACE_STACK_MIN = 8;
bool trace_stack_ok(int entries);
int num_cus();
// Threads per workgroup of the traversal kernel.
constexpr int TRACE_BLOCK_THREADS = 256;
// Rays are K segments (segment k at k * seg); segments [0, nvis) are
// visibility queries toward the distances tl, [nvis, nvis + nseek) BSDF-sampled
// rays that need their hit only when it is an emitting triangle; with
// seek_count (device) the one such segment holds only its first *seek_count
// entries.
// stack_lds: LDS stack entries per lane (8, 16 or 32; 0 = the tuned value);
// prio != 0: the launch's waves take issue priority over concurrent kernels'.
// spill must hold grid * TRACE_BLOCK_THREADS * (STACK_TOTAL - stack_lds) ints.
int trace_rays(const SceneView& sv, const float* ro, const float* rd, int n, const int* count_dev, int K, int seg,
               const float* tl, int nvis, int* tri, float* t, float* u, float* v, int* spill, int grid,
               hipStream_t st, int stack_lds = 0, int prio = 0, int nseek = 0, const int* seek_count = nullptr);
constexpr unsigned long long RAY_BYTES_PER = 28;  // ray buffers: origin, direction, light distance
int trace_grid(int64_t max_rays);
int trace_timing(int enable, double* total_ms, int64_t* launches, int64_t* bytes, double* busy_ms);
int trace_counters(int enable, int64_t* out, int n_out);
size_t spill_bytes(int grid);
int lane_draws(uint32_t lane0, int n_lanes, int n_draws, uint32_t* out, hipStream_t st);
int camera_rays(const CameraK& cam, uint32_t spp, float* o, float* d, int* pix, hipStream_t st);
// The images of n_cams (1..4) cameras of one scene in one wavefront.
int render_forward(psdr_scene* s, const psdr_config& cfg, const psdr_camera* cams, int n_cams, float* const* outs,
                   int keep, hipStream_t st);
// grad_cam: PSDR_CAMERA_GRAD floats or null; cam_acc: camera matrix adjoints
// (cam_grad.h) or null
int render_backward(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam, const float* dL_dI,
                    float* grad_v, float* grad_t, float* grad_e, float* grad_cam, int terms, hipStream_t st);
int interior_backward(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam, const float* dL_dI,
                      float* grad_v, float* grad_t, float* grad_e, float* cam_acc, hipStream_t st);
int fill_f32(float* p, int n, float v, hipStream_t st);
// Edge-sampling boundary terms in two parts: prepare (everything that does not
// depend on dL/dI: hypercube preprocessing, sampler, samples, traversals) and
// finish (the dL/dI-weighted scatter into vertex gradients).
int boundary_primary_prepare(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam, hipStream_t st);
int boundary_primary_finish(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam, const float* dL_dI,
                            float* grad_v, float* cam_acc, hipStream_t st);
int boundary_direct_prepare(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam, hipStream_t st);
int boundary_direct_finish(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam, const float* dL_dI,
                           float* grad_v, hipStream_t st);
// Launches both prepare parts on the scene's auxiliary streams, each waiting
// for the fork event the forward render recorded before its own work.
int boundary_prepare_async(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam);
// Makes `st` wait for everything issued on the auxiliary streams.
int aux_join(psdr_scene* s, hipStream_t st);
}  // namespace psdr
