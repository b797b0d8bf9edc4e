// Scene derivation and LBVH construction on gfx950.
//
//  * k_tri_prep       : per-triangle record (src/mesh.cc:5-67).
//  * k_edge_geom      : per-edge geometry (src/mesh.cc:97-204) over a canonical
//                       (i<j)-sorted edge list built once at scene creation.
//  * k_sampler / k_chunk_* : DiscreteSampler tables (src/sampler.h:86-96) with
//                       the project's fixed fp32 order (chunks of 256, groups of
//                       64 chunk totals, group totals; each level left to right).
//  * LBVH (replaces the OptiX GAS, src/scene_optix.cc:187-238): 30-bit Morton
//    codes of padded triangle-box centres (bit 30 set for triangles whose largest
//    padded extent is at most 1/8 of the centre-bounds extent), a stable LSD radix sort (LDS digit
//    histograms + wave-ballot multisplit ranking), the Karras (2012) hierarchy
//    over keys (morton << 32 | sorted position), and node boxes computed as the
//    union of each internal node's sorted leaf range (chunk boxes, then one wave
//    per node; no inter-workgroup atomics).
//    The integer output is deterministic and matches the CPU restatement exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "scene.h"
#include "traverse.h"

namespace psdr {

namespace {

constexpr int SCAN_CHUNK = 256;
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;

__device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
__device__ __forceinline__ f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }

// --- triangle records ------------------------------------------------------
__global__ void k_tri_prep(const float* __restrict__ V, const int* __restrict__ faces,
                           const float* __restrict__ tri_uv, int n, float4* __restrict__ geom) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const f3 p0 = ld3(V + 3 * faces[3 * f]);
  const f3 p1 = ld3(V + 3 * faces[3 * f + 1]);
  const f3 p2 = ld3(V + 3 * faces[3 * f + 2]);
  const float* uv = tri_uv + 6 * f;
  const f3 a = p2 - p1;
  const f3 b = p0 - p2;
  const f3 c = p1 - p0;
  const f3 nrm = normalize(cross(c, b));
  const float duv1y = uv[3] - uv[1];
  const float duv2y = uv[5] - uv[1];
  const f3 tg = normalize(duv2y * c + duv1y * b);
  const float al = norm(a), bl = norm(b), cl = norm(c);
  const float p = ((al + bl) + cl) / 2.0f;
  const float area = safe_sqrt(((p * (p - al)) * (p - bl)) * (p - cl));
  geom[5 * f + 0] = make_float4(p0.x, p0.y, p0.z, area);
  geom[5 * f + 1] = make_float4(p1.x, p1.y, p1.z, 0.f);
  geom[5 * f + 2] = make_float4(p2.x, p2.y, p2.z, 0.f);
  geom[5 * f + 3] = make_float4(nrm.x, nrm.y, nrm.z, 0.f);
  geom[5 * f + 4] = make_float4(tg.x, tg.y, tg.z, 0.f);
}

// --- edges -------------------------------------------------------------------
__global__ void k_edge_geom(const float* __restrict__ V, const int* __restrict__ topo, int n,
                            float* __restrict__ geom, float* __restrict__ len) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int* t = topo + 6 * e;
  const f3 p0 = ld3(V + 3 * t[0]);
  const f3 p1 = ld3(V + 3 * t[1]);
  const f3 p2l = ld3(V + 3 * t[4]);
  const f3 p2r = ld3(V + 3 * t[5]);
  f3 nl = normalize(cross(p1 - p0, p2l - p1));
  f3 nr = normalize(cross(p2r - p1, p1 - p0));
  const bool shared = t[2] != t[3];
  f3 out = shared ? normalize(nl + nr) : normalize(cross(p1 - p0, nl));
  const f3 cp = normalize((p0 + p1) - (p2l + p2r));
  if (dot(cp, out) < 0.f && shared && norm(nl - nr) > REAL_EPS) {
    nl = -nl;
    nr = -nr;
    out = -out;
  }
  float* g = geom + 16 * e;
  g[0] = p0.x; g[1] = p0.y; g[2] = p0.z;
  g[3] = p1.x; g[4] = p1.y; g[5] = p1.z;
  g[6] = out.x; g[7] = out.y; g[8] = out.z;
  g[9] = nl.x; g[10] = nl.y; g[11] = nl.z;
  g[12] = nr.x; g[13] = nr.y; g[14] = nr.z;
  const float l = norm(p0 - p1);
  g[15] = l;
  len[e] = l;
}

// --- discrete sampler tables ------------------------------------------------
// One block per segment [seg_begin, seg_begin + seg_count) of w.  tmp holds two
// floats per chunk per segment (at 2*seg_begin/SCAN_CHUNK + ...).
// The tables of segment sgi ([b, b + n) of w) by the calling block; every
// thread of the block calls it (it holds barriers).
__device__ void sampler_segment(const float* __restrict__ w, int b, int n, int sgi, float* __restrict__ pmf,
                                float* __restrict__ cmf, float* __restrict__ totals, float* __restrict__ tmp) {
  const int nchunk = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  float* csum = tmp + 2 * ((b + SCAN_CHUNK - 1) / SCAN_CHUNK + sgi);
  float* coff = csum + nchunk;
  __shared__ float s_total;
  for (int c = threadIdx.x; c < nchunk; c += blockDim.x) {
    float s = 0.f;
    const int e = min(n, (c + 1) * SCAN_CHUNK);
    for (int i = c * SCAN_CHUNK; i < e; ++i) s = s + w[b + i];
    csum[c] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float run = 0.f;
    for (int g0 = 0; g0 < nchunk; g0 += 64) {
      float gr = 0.f;
      for (int c = g0; c < min(nchunk, g0 + 64); ++c) gr = gr + csum[c];
      run = run + gr;
    }
    s_total = run;
    if (totals) totals[sgi] = run;
  }
  __syncthreads();
  const float total = s_total;
  for (int c = threadIdx.x; c < nchunk; c += blockDim.x) {
    float s = 0.f;
    const int e = min(n, (c + 1) * SCAN_CHUNK);
    for (int i = c * SCAN_CHUNK; i < e; ++i) {
      const float p = w[b + i] / total;
      pmf[b + i] = p;
      s = s + p;
    }
    csum[c] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float run = 0.f;
    for (int g0 = 0; g0 < nchunk; g0 += 64) {
      float gr = 0.f;
      for (int c = g0; c < min(nchunk, g0 + 64); ++c) {
        coff[c] = run + gr;
        gr = gr + csum[c];
      }
      run = run + gr;
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nchunk; c += blockDim.x) {
    float s = 0.f;
    const float off = coff[c];
    const int e = min(n, (c + 1) * SCAN_CHUNK);
    for (int i = c * SCAN_CHUNK; i < e; ++i) {
      s = s + pmf[b + i];
      cmf[b + i] = off + s;
    }
  }
  __syncthreads();  // the caller may reuse tmp and s_total for another segment
}
__global__ void k_sampler(const float* __restrict__ w, const int* __restrict__ seg_begin,
                          const int* __restrict__ seg_count, int n_static, float* __restrict__ pmf,
                          float* __restrict__ cmf, float* __restrict__ totals, float* __restrict__ tmp) {
  const int sgi = blockIdx.x;
  sampler_segment(w, seg_begin ? seg_begin[sgi] : 0, seg_count ? seg_count[sgi] : n_static, sgi, pmf, cmf, totals,
                  tmp);
}

// Three-level fixed order for long segments: chunks of 256 (sequential), groups
// of 64 chunk totals (sequential), group totals (sequential).  Chunks are staged
// through LDS with coalesced loads; one thread per chunk adds in order.
constexpr int SCAN_GROUP = 64;
constexpr int CHUNKS_PER_BLOCK = 4;
// csum[c] = sum of chunk c of (w / *div if div else w); writes pmf when div
__global__ __launch_bounds__(256) void k_chunk_sums(const float* __restrict__ w, const float* __restrict__ div,
                                                     int n, float* __restrict__ csum, float* __restrict__ pmf_out) {
  __shared__ float tile[CHUNKS_PER_BLOCK * SCAN_CHUNK];
  const int c0 = blockIdx.x * CHUNKS_PER_BLOCK;
  const int base = c0 * SCAN_CHUNK;
  const int m = min(n - base, CHUNKS_PER_BLOCK * SCAN_CHUNK);
  if (m <= 0) return;
  const float total = div ? *div : 1.f;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    float v = w[base + i];
    if (div) {
      v = v / total;
      pmf_out[base + i] = v;
    }
    tile[i] = v;
  }
  __syncthreads();
  const int c = threadIdx.x;
  if (c < CHUNKS_PER_BLOCK && c * SCAN_CHUNK < m) {
    const int e = min(m, (c + 1) * SCAN_CHUNK);
    float s = 0.f;
    for (int i = c * SCAN_CHUNK; i < e; ++i) s = s + tile[i];
    csum[c0 + c] = s;
  }
}
// per group: exclusive scan of its chunk totals in place, group total -> gsum
__global__ void k_group_scan(float* __restrict__ csum, int nchunk, float* __restrict__ gsum) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int ngroup = (nchunk + SCAN_GROUP - 1) / SCAN_GROUP;
  if (g >= ngroup) return;
  float run = 0.f;
  const int e = min(nchunk, (g + 1) * SCAN_GROUP);
  for (int c = g * SCAN_GROUP; c < e; ++c) {
    const float x = csum[c];
    csum[c] = run;
    run = run + x;
  }
  gsum[g] = run;
}
// exclusive scan of group totals in place; total -> *total
__global__ void k_top_scan(float* __restrict__ gsum, int ngroup, float* __restrict__ total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float run = 0.f;
  for (int g = 0; g < ngroup; ++g) {
    const float x = gsum[g];
    gsum[g] = run;
    run = run + x;
  }
  *total = run;
}
__global__ __launch_bounds__(256) void k_chunk_cmf(const float* __restrict__ pmf, int n,
                                                    const float* __restrict__ inner, const float* __restrict__ goff,
                                                    float* __restrict__ cmf) {
  __shared__ float tile[CHUNKS_PER_BLOCK * SCAN_CHUNK];
  const int c0 = blockIdx.x * CHUNKS_PER_BLOCK;
  const int base = c0 * SCAN_CHUNK;
  const int m = min(n - base, CHUNKS_PER_BLOCK * SCAN_CHUNK);
  if (m <= 0) return;
  for (int i = threadIdx.x; i < m; i += blockDim.x) tile[i] = pmf[base + i];
  __syncthreads();
  const int c = threadIdx.x;
  if (c < CHUNKS_PER_BLOCK && c * SCAN_CHUNK < m) {
    const float off = goff[(c0 + c) / SCAN_GROUP] + inner[c0 + c];
    const int e = min(m, (c + 1) * SCAN_CHUNK);
    float s = 0.f;
    for (int i = c * SCAN_CHUNK; i < e; ++i) {
      s = s + tile[i];
      tile[i] = off + s;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += blockDim.x) cmf[base + i] = tile[i];
}

// light-triangle areas gathered contiguously per light, then per-light sampler.
__global__ void k_light_areas(const float4* __restrict__ geom, const LightK* __restrict__ lights, int n_lights,
                              float* __restrict__ out, int* __restrict__ seg_begin, int* __restrict__ seg_count) {
  const int l = blockIdx.x;
  if (l >= n_lights) return;
  const LightK L = lights[l];
  for (int k = threadIdx.x; k < L.tri_count; k += blockDim.x) out[L.pmf_off + k] = geom[5 * (L.tri_begin + k)].w;
  if (threadIdx.x == 0) {
    seg_begin[l] = L.pmf_off;
    seg_count[l] = L.tri_count;
  }
}

// AreaLight power (src/emitter.cc:10-11), the environment lights' fixed 4*pi
// (src/emitter.cc:129, 247) and the per-object light weights.
__global__ void k_light_power(const ObjectK* __restrict__ objs, int n_objects, const float* __restrict__ emit,
                              const float* __restrict__ light_area, float* __restrict__ light_power,
                              float* __restrict__ obj_w) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_objects) return;
  const int l = objs[o].light;
  float w = 0.f;
  if (l >= 0) {
    if (objs[o].mat == PSDR_MAT_ENV_MAP || objs[o].mat == PSDR_MAT_ENV_SH9) {
      w = REAL_4PI;
    } else {
      const float lum = (emit[3 * l] * 0.2126f + emit[3 * l + 1] * 0.7152f) + emit[3 * l + 2] * 0.0722f;
      w = (REAL_2PI * lum) * light_area[l];
    }
    light_power[l] = w;
  }
  obj_w[o] = w;
}

// The light tables of a scene with few lights in one workgroup, in the order
// of the separate launches (k_light_areas, per-light k_sampler, k_light_power,
// the object k_sampler): same values, four launches fewer on the path every
// renderD step starts with.  w: light-triangle areas, then object weights.
__global__ void k_light_tables(const float4* __restrict__ geom, const LightK* __restrict__ lights, int n_lights,
                               const ObjectK* __restrict__ objs, int n_objects, const float* __restrict__ emit,
                               float* __restrict__ w, float* __restrict__ ltri_pmf, float* __restrict__ ltri_cmf,
                               float* __restrict__ light_area, float* __restrict__ light_power,
                               float* __restrict__ obj_pmf, float* __restrict__ obj_cmf,
                               float* __restrict__ lights_total, float* __restrict__ tmp) {
  for (int l = 0; l < n_lights; ++l) {
    const LightK L = lights[l];
    for (int k = threadIdx.x; k < L.tri_count; k += blockDim.x) w[L.pmf_off + k] = geom[5 * (L.tri_begin + k)].w;
  }
  __syncthreads();
  for (int l = 0; l < n_lights; ++l) {
    const LightK L = lights[l];
    sampler_segment(w, L.pmf_off, L.tri_count, l, ltri_pmf, ltri_cmf, light_area, tmp);
  }
  for (int o = threadIdx.x; o < n_objects; o += blockDim.x) {
    const int l = objs[o].light;
    float v = 0.f;
    if (l >= 0) {
      if (objs[o].mat == PSDR_MAT_ENV_MAP || objs[o].mat == PSDR_MAT_ENV_SH9) {
        v = REAL_4PI;
      } else {
        const float lum = (emit[3 * l] * 0.2126f + emit[3 * l + 1] * 0.7152f) + emit[3 * l + 2] * 0.0722f;
        v = (REAL_2PI * lum) * light_area[l];
      }
      light_power[l] = v;
    }
    w[o] = v;
  }
  __syncthreads();
  sampler_segment(w, 0, n_objects, 0, obj_pmf, obj_cmf, lights_total, tmp);
}

// --- LBVH --------------------------------------------------------------------
__device__ __forceinline__ void tri_box(const float4* geom, int f, float lo[3], float hi[3]) {
  const float4 a = geom[5 * f], b = geom[5 * f + 1], c = geom[5 * f + 2];
  const float xs[3][3] = {{a.x, b.x, c.x}, {a.y, b.y, c.y}, {a.z, b.z, c.z}};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float l = fminf(xs[k][0], fminf(xs[k][1], xs[k][2]));
    const float h = fmaxf(xs[k][0], fmaxf(xs[k][1], xs[k][2]));
    const float pad = (fabsf(l) + fabsf(h)) * 1e-6f + 1e-30f;
    lo[k] = l - pad;
    hi[k] = h + pad;
  }
}
__device__ __forceinline__ void tri_center(const float4* geom, int f, float c[3]) {
  const float4 a = geom[5 * f], b = geom[5 * f + 1], d = geom[5 * f + 2];
  const float xs[3][3] = {{a.x, b.x, d.x}, {a.y, b.y, d.y}, {a.z, b.z, d.z}};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float l = fminf(xs[k][0], fminf(xs[k][1], xs[k][2]));
    const float h = fmaxf(xs[k][0], fmaxf(xs[k][1], xs[k][2]));
    c[k] = (l + h) * 0.5f;
  }
}

// centre bounds, phase 1: per-block min/max -> part[6*block]
__global__ void k_center_bounds(const float4* __restrict__ geom, int n, float* __restrict__ part) {
  __shared__ float s[6][256];
  float mn[3] = {F_INF, F_INF, F_INF}, mx[3] = {-F_INF, -F_INF, -F_INF};
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    float c[3];
    tri_center(geom, f, c);
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], c[k]);
      mx[k] = fmaxf(mx[k], c[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    s[k][threadIdx.x] = mn[k];
    s[3 + k][threadIdx.x] = mx[k];
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + st]);
        s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}
// phase 2 (one block): reduce parts -> bounds[6]
__global__ void k_center_bounds2(const float* __restrict__ part, int nparts, float* __restrict__ bounds) {
  if (threadIdx.x >= 6) return;
  const int k = threadIdx.x;
  float v = k < 3 ? F_INF : -F_INF;
  for (int p = 0; p < nparts; ++p) v = k < 3 ? fminf(v, part[6 * p + k]) : fmaxf(v, part[6 * p + k]);
  bounds[k] = v;
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ void morton_one(const float4* __restrict__ geom, int f, const float* bounds,
                                           uint32_t* __restrict__ morton, uint32_t* __restrict__ keys,
                                           uint32_t* __restrict__ vals) {
  float c[3];
  tri_center(geom, f, c);
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float ext = bounds[3 + k] - bounds[k];
    const float x = ext > 0.f ? (c[k] - bounds[k]) / ext : 0.5f;
    q[k] = static_cast<uint32_t>(fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f));
  }
  uint32_t m = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
  // size partition: small triangles get bit 30, so the root separates the few
  // large ones (ground planes, lights) from the rest instead of letting their
  // boxes inflate every ancestor of a deep leaf
  float lo[3], hi[3], cmax = 0.f, emax = 0.f;
  tri_box(geom, f, lo, hi);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    emax = fmaxf(emax, bounds[3 + k] - bounds[k]);
    cmax = fmaxf(cmax, hi[k] - lo[k]);
  }
  if (!(cmax > 0.125f * emax)) m |= 1u << 30;
  morton[f] = m;
  keys[f] = m;
  vals[f] = static_cast<uint32_t>(f);
}
__global__ void k_morton(const float4* __restrict__ geom, int n, const float* __restrict__ bounds,
                         uint32_t* __restrict__ morton, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) morton_one(geom, f, bounds, morton, keys, vals);
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                            uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const int i = blockIdx.x * SORT_TILE + r * SORT_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of m uint32 values, one block of 1024 threads
__global__ __launch_bounds__(1024) void k_scan_u32(uint32_t* __restrict__ a, int m) {
  __shared__ uint32_t s[1024];
  const int per = (m + 1023) / 1024;
  const int b = threadIdx.x * per, e = min(m, b + per);
  uint32_t sum = 0;
  for (int i = b; i < e; ++i) sum += a[i];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - sum;
  for (int i = b; i < e; ++i) {
    const uint32_t x = a[i];
    a[i] = run;
    run += x;
  }
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin,
                                                               uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout, int n, int shift,
                                                               const uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_cnt[SORT_THREADS / 64][256];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  s_base[tid] = hist[tid * gridDim.x + blockIdx.x];
  for (int w = 0; w < SORT_THREADS / 64; ++w) s_cnt[w][tid] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const int i = blockIdx.x * SORT_TILE + r * SORT_THREADS + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t digit = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool on = (digit >> bit) & 1u;
      const uint64_t bb = __ballot(on);
      peers &= on ? bb : ~bb;
    }
    const uint32_t rank = __popcll(peers & lt);
    if (valid && rank == 0) s_cnt[wave][digit] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t off = s_base[digit];
      for (int w = 0; w < wave; ++w) off += s_cnt[w][digit];
      kout[off + rank] = key;
      vout[off + rank] = val;
    }
    __syncthreads();
    uint32_t tot = 0;
    for (int w = 0; w < SORT_THREADS / 64; ++w) {
      tot += s_cnt[w][tid];
      s_cnt[w][tid] = 0;
    }
    s_base[tid] += tot;
    __syncthreads();
  }
}

__device__ __forceinline__ int kdelta(const uint32_t* ms, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = (static_cast<uint64_t>(ms[i]) << 32) | static_cast<uint32_t>(i);
  const uint64_t b = (static_cast<uint64_t>(ms[j]) << 32) | static_cast<uint32_t>(j);
  return __clzll(static_cast<long long>(a ^ b));
}

__global__ void k_karras(const uint32_t* __restrict__ ms, int n, int* __restrict__ child, int* __restrict__ parent,
                         int2* __restrict__ range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (kdelta(ms, n, i, i + 1) - kdelta(ms, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = kdelta(ms, n, i, i - d);
  int lmax = 2;
  while (kdelta(ms, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(ms, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(ms, n, i, j);
  int s = 0, div = 2, t;
  do {
    t = (l + div - 1) / div;
    if (kdelta(ms, n, i, i + (s + t) * d) > dnode) s += t;
    div *= 2;
  } while (t > 1);
  const int g = i + s * d + min(d, 0);
  const int left = (min(i, j) == g) ? (g | static_cast<int>(0x80000000u)) : g;
  const int right = (max(i, j) == g + 1) ? ((g + 1) | static_cast<int>(0x80000000u)) : g + 1;
  child[2 * i] = left;
  child[2 * i + 1] = right;
  range[i] = make_int2(min(i, j), max(i, j));
  parent[left < 0 ? (n - 1 + (left & 0x7fffffff)) : left] = i;
  parent[right < 0 ? (n - 1 + (right & 0x7fffffff)) : right] = i;
  if (i == 0) parent[0] = -1;
}

// Leaf records in sorted order, the sorted order itself and each leaf's padded box.
// Triangle ids are first + f for records geom[f].
__global__ void k_leaf(const float4* __restrict__ geom, const uint32_t* __restrict__ sorted, int n, int first,
                       float4* __restrict__ leaf, uint32_t* __restrict__ order, float* __restrict__ lbox) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int f = sorted ? static_cast<int>(sorted[k]) : k;
  if (order) order[k] = static_cast<uint32_t>(f);
  if (lbox) {
    float lo[3], hi[3];
    tri_box(geom, f, lo, hi);
    for (int a = 0; a < 3; ++a) {
      lbox[6 * k + a] = lo[a];
      lbox[6 * k + 3 + a] = hi[a];
    }
  }
  const float4 a = geom[5 * f], b = geom[5 * f + 1], c = geom[5 * f + 2];
  const f3 p0 = xyz(a), p1 = xyz(b), p2 = xyz(c);
  const f3 e1 = p1 - p0, e2 = p2 - p0;
  leaf[3 * k] = make_float4(p0.x, p0.y, p0.z, __int_as_float(first + f));
  leaf[3 * k + 1] = make_float4(e1.x, e1.y, e1.z, 0.f);
  leaf[3 * k + 2] = make_float4(e2.x, e2.y, e2.z, 0.f);
}

// Boxes of internal nodes without a bottom-up pass: internal node i covers
// the sorted leaves range[i] = [first, last], and its box is the union of those
// leaves' boxes.  min/max are exact, so any grouping gives the same bits.  A
// wave reduces a node's range from 64-leaf chunk boxes plus the partial chunks
// at its ends; no inter-workgroup communication is needed.
__device__ __forceinline__ void box_acc(float lo[3], float hi[3], const float* b) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = fminf(lo[a], b[a]);
    hi[a] = fmaxf(hi[a], b[3 + a]);
  }
}
__device__ __forceinline__ void wave_box_reduce(float lo[3], float hi[3]) {
  for (int m = 32; m > 0; m >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], m, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], m, 64));
    }
}

// one wave per 64-leaf chunk
__global__ __launch_bounds__(256) void k_chunk_boxes(const float* __restrict__ lbox, int n, float* __restrict__ cbox) {
  const int c = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = static_cast<int>(threadIdx.x & 63);
  if (c * 64 >= n) return;  // whole wave leaves together
  float lo[3] = {F_INF, F_INF, F_INF}, hi[3] = {-F_INF, -F_INF, -F_INF};
  const int k = c * 64 + lane;
  if (k < n) box_acc(lo, hi, lbox + 6 * k);
  wave_box_reduce(lo, hi);
  if (lane == 0)
    for (int a = 0; a < 3; ++a) {
      cbox[6 * c + a] = lo[a];
      cbox[6 * c + 3 + a] = hi[a];
    }
}

// one wave per internal node
__global__ __launch_bounds__(256) void k_node_boxes(const float* __restrict__ lbox, const float* __restrict__ cbox,
                                                    const int2* __restrict__ range, int n, float* __restrict__ nbox) {
  const int i = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = static_cast<int>(threadIdx.x & 63);
  if (i >= n - 1) return;  // whole wave leaves together
  const int2 r = range[i];
  const int first = r.x, last = r.y;
  float lo[3] = {F_INF, F_INF, F_INF}, hi[3] = {-F_INF, -F_INF, -F_INF};
  const int c0 = (first + 63) >> 6, c1 = (last + 1) >> 6;  // full chunks [c0, c1)
  if (c0 >= c1) {
    for (int k = first + lane; k <= last; k += 64) box_acc(lo, hi, lbox + 6 * k);
  } else {
    for (int k = first + lane; k < 64 * c0; k += 64) box_acc(lo, hi, lbox + 6 * k);
    for (int c = c0 + lane; c < c1; c += 64) box_acc(lo, hi, cbox + 6 * c);
    for (int k = 64 * c1 + lane; k <= last; k += 64) box_acc(lo, hi, lbox + 6 * k);
  }
  wave_box_reduce(lo, hi);
  if (lane == 0)
    for (int a = 0; a < 3; ++a) {
      nbox[6 * i + a] = lo[a];
      nbox[6 * i + 3 + a] = hi[a];
    }
}

// node i = the 48-byte traversal node of its two child boxes (pack_node);
// box keeps the exact 12 planes
__global__ void k_pack_nodes(const float* __restrict__ lbox, const float* __restrict__ nbox,
                             const int* __restrict__ child, int n, float* __restrict__ box,
                             float4* __restrict__ node) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  float b[12];
  for (int s = 0; s < 2; ++s) {
    const int c = child[2 * i + s];
    const float* src = c < 0 ? lbox + 6 * (c & 0x7fffffff) : nbox + 6 * c;
    for (int a = 0; a < 6; ++a) b[6 * s + a] = src[a];
  }
  for (int a = 0; a < 12; ++a) box[12 * i + a] = b[a];
  pack_node(b, b + 6, child[2 * i], child[2 * i + 1], node + 3 * i);
}

inline int blocks(long long n, int t) { return static_cast<int>((n + t - 1) / t); }

}  // namespace

// Sampler tables over `nseg` segments (device arrays seg_begin/seg_count).
static int launch_sampler(const float* w, const int* seg_begin, const int* seg_count, int nseg, float* pmf,
                          float* cmf, float* totals, float* tmp, hipStream_t st) {
  if (nseg <= 0) return PSDR_OK;
  hipLaunchKernelGGL(k_sampler, dim3(nseg), dim3(256), 0, st, w, seg_begin, seg_count, 0, pmf, cmf, totals, tmp);
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

int sampler_tables(const float* w, int n, float* pmf, float* cmf, float* total, float* tmp, hipStream_t st) {
  if (n <= 0) return PSDR_OK;
  const int nchunk = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  if (nchunk <= 8) {
    hipLaunchKernelGGL(k_sampler, dim3(1), dim3(256), 0, st, w, nullptr, nullptr, n, pmf, cmf, total, tmp);
    PSDR_HIP_CHECK(hipGetLastError());
    return PSDR_OK;
  }
  // tmp: [inner nchunk][gsum ngroup][tot 2]
  const int ngroup = (nchunk + SCAN_GROUP - 1) / SCAN_GROUP;
  float* inner = tmp;
  float* gsum = tmp + nchunk;
  float* tot = gsum + ngroup;
  const int nb = blocks(nchunk, CHUNKS_PER_BLOCK);
  const int ng = blocks(ngroup, 64);
  hipLaunchKernelGGL(k_chunk_sums, dim3(nb), dim3(256), 0, st, w, nullptr, n, inner, nullptr);
  hipLaunchKernelGGL(k_group_scan, dim3(ng), dim3(64), 0, st, inner, nchunk, gsum);
  hipLaunchKernelGGL(k_top_scan, dim3(1), dim3(64), 0, st, gsum, ngroup, tot);
  hipLaunchKernelGGL(k_chunk_sums, dim3(nb), dim3(256), 0, st, w, tot, n, inner, pmf);
  hipLaunchKernelGGL(k_group_scan, dim3(ng), dim3(64), 0, st, inner, nchunk, gsum);
  hipLaunchKernelGGL(k_top_scan, dim3(1), dim3(64), 0, st, gsum, ngroup, tot + 1);
  hipLaunchKernelGGL(k_chunk_cmf, dim3(nb), dim3(256), 0, st, pmf, n, inner, gsum, cmf);
  PSDR_HIP_CHECK(hipGetLastError());
  if (total) PSDR_HIP_CHECK(hipMemcpyAsync(total, tot, 4, hipMemcpyDeviceToDevice, st));
  return PSDR_OK;
}

int scene_derive(psdr_scene* s, hipStream_t st) {
  const int nt = s->n_tris;
  if (nt > 0) {
    hipLaunchKernelGGL(k_tri_prep, dim3(blocks(nt, 256)), dim3(256), 0, st, s->vertices, s->faces.as<int>(),
                       s->tri_uv.as<float>(), nt, s->tri_geom.as<float4>());
    PSDR_HIP_CHECK(hipGetLastError());
  }
  // scratch layout in scan_tmp: [ints: seg_begin(nl), seg_count(nl), one(1)] [floats...]
  const int nl = s->n_lights;
  const int ne = s->n_edges;
  const int nobj = s->n_objects;
  int* segi = s->scan_tmp.as<int>();
  float* wbuf = reinterpret_cast<float*>(segi + 2 * nl + 4);
  const size_t wmax = std::max<size_t>({static_cast<size_t>(s->light_tri_total), static_cast<size_t>(ne),
                                        static_cast<size_t>(nobj), 1});
  float* tmp = wbuf + wmax;
  // few lights: the light tables in one workgroup
  const bool light_fused = nl > 0 && nl <= 16 && s->light_tri_total <= 16384 && nobj <= 16384;
  if (light_fused) {
    hipLaunchKernelGGL(k_light_tables, dim3(1), dim3(256), 0, st, s->tri_geom.as<float4>(), s->lights.as<LightK>(),
                       nl, s->objs_dev, nobj, s->light_emit, wbuf, s->light_tri_pmf.as<float>(),
                       s->light_tri_cmf.as<float>(), s->light_area.as<float>(), s->light_power.as<float>(),
                       s->obj_pmf.as<float>(), s->obj_cmf.as<float>(), s->lights_total.as<float>(), tmp);
    PSDR_HIP_CHECK(hipGetLastError());
  }
  // lights: areas -> per-light tri samplers and total areas
  if (nl > 0 && !light_fused) {
    hipLaunchKernelGGL(k_light_areas, dim3(nl), dim3(256), 0, st, s->tri_geom.as<float4>(),
                       s->lights.as<LightK>(), nl, wbuf, segi, segi + nl);
    PSDR_HIP_CHECK(hipGetLastError());
    int rc = launch_sampler(wbuf, segi, segi + nl, nl, s->light_tri_pmf.as<float>(), s->light_tri_cmf.as<float>(),
                            s->light_area.as<float>(), tmp, st);
    if (rc) return rc;
  }
  // light power + light pick over objects
  if (nobj > 0 && !light_fused) {
    hipLaunchKernelGGL(k_light_power, dim3(blocks(nobj, 64)), dim3(64), 0, st, s->objs_dev, nobj,
                       s->light_emit, s->light_area.as<float>(), s->light_power.as<float>(), wbuf);
    PSDR_HIP_CHECK(hipGetLastError());
    int* cnt = segi + 2 * nl;  // holds nobj (uploaded at creation)
    int rc = launch_sampler(wbuf, nullptr, cnt, 1, s->obj_pmf.as<float>(), s->obj_cmf.as<float>(),
                            s->lights_total.as<float>(), tmp, st);
    if (rc) return rc;
  }
  // edges
  if (ne > 0) {
    hipLaunchKernelGGL(k_edge_geom, dim3(blocks(ne, 256)), dim3(256), 0, st, s->vertices, s->edge_topo.as<int>(),
                       ne, s->edge_geom.as<float>(), wbuf);
    PSDR_HIP_CHECK(hipGetLastError());
    int rc = sampler_tables(wbuf, ne, s->edge_pmf.as<float>(), s->edge_cmf.as<float>(), nullptr, tmp, st);
    if (rc) return rc;
  }
  return build_lbvh(s, st);
}

// exclusive scan of m uint32 in three launches: per-block sums, a single-block
// scan of those, per-block local scans plus the block offset
constexpr int XS_THREADS = 256, XS_PER = 8, XS_BLOCK = XS_THREADS * XS_PER;

__global__ __launch_bounds__(XS_THREADS) void k_xscan_sums(const uint32_t* __restrict__ a, int m,
                                                           uint32_t* __restrict__ part) {
  __shared__ uint32_t red[XS_THREADS / 64];
  const int base = blockIdx.x * XS_BLOCK;
  uint32_t sum = 0;
  for (int j = 0; j < XS_PER; ++j) {
    const int i = base + j * XS_THREADS + static_cast<int>(threadIdx.x);
    if (i < m) sum += a[i];
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < XS_THREADS / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(XS_THREADS) void k_xscan_down(uint32_t* __restrict__ a, int m,
                                                           const uint32_t* __restrict__ part) {
  __shared__ uint32_t s[XS_THREADS];
  const int base = blockIdx.x * XS_BLOCK + static_cast<int>(threadIdx.x) * XS_PER;
  uint32_t v[XS_PER];
  uint32_t sum = 0;
  for (int j = 0; j < XS_PER; ++j) {
    v[j] = base + j < m ? a[base + j] : 0u;
    sum += v[j];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < XS_THREADS; off <<= 1) {
    const uint32_t x = threadIdx.x >= static_cast<unsigned>(off) ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = part[blockIdx.x] + s[threadIdx.x] - sum;
  for (int j = 0; j < XS_PER; ++j) {
    if (base + j < m) a[base + j] = run;
    run += v[j];
  }
}

size_t radix_hist_words(int n) {
  const size_t m = 256ull * std::max(blocks(n, SORT_TILE), 1);
  return m + (m + XS_BLOCK - 1) / XS_BLOCK;
}

int exclusive_scan_u32(uint32_t* a, int m, uint32_t* part, hipStream_t st) {
  const int nb = blocks(m, XS_BLOCK);
  if (nb <= 16) {  // up to 32k words: one block is cheaper than three launches
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, st, a, m);
  } else {
    hipLaunchKernelGGL(k_xscan_sums, dim3(nb), dim3(XS_THREADS), 0, st, a, m, part);
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, st, part, nb);
    hipLaunchKernelGGL(k_xscan_down, dim3(nb), dim3(XS_THREADS), 0, st, a, m, part);
  }
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* hist, int n, int passes,
                     hipStream_t st) {
  if (n <= 0) return PSDR_OK;
  const int tiles = blocks(n, SORT_TILE);
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = 8 * pass;
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(SORT_THREADS), 0, st, k0, n, shift, hist);
    const int rc = exclusive_scan_u32(hist, 256 * tiles, hist + 256 * tiles, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(SORT_THREADS), 0, st, k0, v0, k1, v1, n, shift, hist);
    PSDR_HIP_CHECK(hipGetLastError());
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  return PSDR_OK;
}

int build_lbvh(psdr_scene* s, hipStream_t st) {
  const int n = s->n_bvh;
  if (s->n_seek > 0) {
    // the emitting triangles' leaf records again, in id order (seek leaves)
    int off = 0;
    for (const auto& r : s->seek_ranges) {
      hipLaunchKernelGGL(k_leaf, dim3(1), dim3(64), 0, st, s->tri_geom.as<float4>() + 5ll * r.first, nullptr,
                         r.second, r.first, s->seek_leaf.as<float4>() + 3ll * off, nullptr, nullptr);
      PSDR_HIP_CHECK(hipGetLastError());
      off += r.second;
    }
  }
  if (s->n_tris > n) {
    // leaves tested by every ray (an environment cube), after the tree's own
    hipLaunchKernelGGL(k_leaf, dim3(1), dim3(64), 0, st, s->tri_geom.as<float4>() + 5ll * n, nullptr, s->n_tris - n,
                       n, s->leaf.as<float4>() + 3ll * n, nullptr, nullptr);
    PSDR_HIP_CHECK(hipGetLastError());
  }
  if (n < 2) {
    if (n == 1) {
      hipLaunchKernelGGL(k_leaf, dim3(1), dim3(64), 0, st, s->tri_geom.as<float4>(), nullptr, 1, 0,
                         s->leaf.as<float4>(), nullptr, nullptr);
      PSDR_HIP_CHECK(hipGetLastError());
    }
    return PSDR_OK;
  }
  uint32_t* k0 = s->sort_keys.as<uint32_t>();
  uint32_t* k1 = k0 + n;
  uint32_t* v0 = s->sort_vals.as<uint32_t>();
  uint32_t* v1 = v0 + n;
  const int nb = std::min(blocks(n, 256), 1024);
  float* part = s->bounds.as<float>() + 8;
  hipLaunchKernelGGL(k_center_bounds, dim3(nb), dim3(256), 0, st, s->tri_geom.as<float4>(), n, part);
  hipLaunchKernelGGL(k_center_bounds2, dim3(1), dim3(64), 0, st, part, nb, s->bounds.as<float>());
  hipLaunchKernelGGL(k_morton, dim3(blocks(n, 256)), dim3(256), 0, st, s->tri_geom.as<float4>(), n,
                     s->bounds.as<float>(), s->morton.as<uint32_t>(), k0, v0);
  PSDR_HIP_CHECK(hipGetLastError());
  const int rc = radix_sort_pairs(k0, v0, k1, v1, s->sort_hist.as<uint32_t>(), n, 4, st);
  if (rc) return rc;
  // after an even number of passes the sorted data is back in k0/v0 (= buffer start)
  float* lbox = s->bvh_boxes.as<float>();
  float* nbox = lbox + 6ull * n;
  float* cbox = nbox + 6ull * (n - 1);
  int2* range = s->ranges.as<int2>();
  hipLaunchKernelGGL(k_karras, dim3(blocks(n - 1, 256)), dim3(256), 0, st, k0, n, s->child.as<int>(),
                     s->parent.as<int>(), range);
  hipLaunchKernelGGL(k_leaf, dim3(blocks(n, 256)), dim3(256), 0, st, s->tri_geom.as<float4>(), v0, n, 0,
                     s->leaf.as<float4>(), s->order.as<uint32_t>(), lbox);
  const int chunks = (n + 63) / 64;
  hipLaunchKernelGGL(k_chunk_boxes, dim3(blocks(64ll * chunks, 256)), dim3(256), 0, st, lbox, n, cbox);
  hipLaunchKernelGGL(k_node_boxes, dim3(blocks(64ll * (n - 1), 256)), dim3(256), 0, st, lbox, cbox, range, n, nbox);
  hipLaunchKernelGGL(k_pack_nodes, dim3(blocks(n - 1, 256)), dim3(256), 0, st, lbox, nbox, s->child.as<int>(), n,
                     s->box.as<float>(), s->node.as<float4>());
  PSDR_HIP_CHECK(hipGetLastError());
  return PSDR_OK;
}

}  // namespace psdr
