// Workgroup-level aggregation of scatter-add gradients in LDS.
//
// Gradient scatters of neighbouring samples land on the same few addresses
// (the 16 samples of a pixel hit the same triangle; boundary samples sorted by
// pixel or edge position hit the same edges), and global float atomics on one
// address serialise in L2.  A block first adds its contributions into a small
// open-addressing table in LDS keyed by the destination triple (3 consecutive
// floats: a vertex or an RGB texel), then adds each table entry to global
// memory once.  Float sums are order-dependent, so results agree with direct
// atomics to rounding, as any atomic accumulation does.
#pragma once

#include "common.h"

namespace psdr {

template <int SLOTS>
struct LdsGradCache {
  static_assert((SLOTS & (SLOTS - 1)) == 0, "SLOTS must be a power of two");
  static constexpr int PROBES = 8;
  int* keys;    // [SLOTS], -1 = empty; key = (array << 30) | float offset of the triple
  float* vals;  // [SLOTS][3]

  // all threads of the block; followed by __syncthreads() before use
  __device__ void init() {
    for (int s = threadIdx.x; s < SLOTS; s += blockDim.x) {
      keys[s] = -1;
      vals[3 * s] = 0.f;
      vals[3 * s + 1] = 0.f;
      vals[3 * s + 2] = 0.f;
    }
  }

  // g into dst[off .. off+2] of array `arr` (0 or 1, base given at flush)
  __device__ void add(float* base, int arr, int64_t off, f3 g) {
    if (g.x == 0.f && g.y == 0.f && g.z == 0.f) return;
    if (off < (1ll << 30)) {
      const int key = (arr << 30) | static_cast<int>(off);
      uint32_t h = (static_cast<uint32_t>(key) * 2654435761u) >> 16;
      for (int p = 0; p < PROBES; ++p, ++h) {
        const int s = static_cast<int>(h & (SLOTS - 1));
        const int prev = atomicCAS(&keys[s], -1, key);
        if (prev == -1 || prev == key) {
          atomicAdd(&vals[3 * s], g.x);
          atomicAdd(&vals[3 * s + 1], g.y);
          atomicAdd(&vals[3 * s + 2], g.z);
          return;
        }
      }
    }
    float* d = base + off;
    if (g.x != 0.f) atomicAdd(d, g.x);
    if (g.y != 0.f) atomicAdd(d + 1, g.y);
    if (g.z != 0.f) atomicAdd(d + 2, g.z);
  }

  // all threads of the block, after __syncthreads()
  __device__ void flush(float* base0, float* base1) {
    for (int s = threadIdx.x; s < SLOTS; s += blockDim.x) {
      const int key = keys[s];
      if (key < 0) continue;
      float* d = ((key >> 30) ? base1 : base0) + (key & ((1 << 30) - 1));
      const float x = vals[3 * s], y = vals[3 * s + 1], z = vals[3 * s + 2];
      if (x != 0.f) atomicAdd(d, x);
      if (y != 0.f) atomicAdd(d + 1, y);
      if (z != 0.f) atomicAdd(d + 2, z);
    }
  }
};

}  // namespace psdr
