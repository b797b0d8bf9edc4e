// Closest-hit LBVH traversal for gfx950 (replaces optixTrace with the
// raygen/closesthit/miss programs of src/cuda/device_programs.cu:13-47).
//
// BVH2 with both child boxes stored in the parent: a 48-byte node (three
// 16-byte loads) holds the two boxes' 12 planes, and the child ids ride in the
// low mantissa bytes of four of them (node_children); leaves are single
// triangles stored in LBVH order as {p0, id}{e1}{e2}.  The traversal stack lives in LDS (STACK_LDS entries per
// lane, laid out [entry][lane] so a wave's pushes hit distinct banks) and
// overflows into a per-thread global area.  The tree is a Karras hierarchy over
// 64-bit keys (31-bit Morton code << 32 | sorted position; the top bit is always
// zero): an internal node's common-prefix length lies in [1, 63] and grows
// strictly from parent to child, so a root-to-leaf path has at most 63 internal
// nodes, and the stack, which holds at most one entry per level, never needs
// more than the 64 entries it has.
//
// Result contract: smallest t >= 1e-3 (Moller-Trumbore, fixed op order); an
// exactly equal t keeps the smaller global triangle id, so the answer does not
// depend on traversal order.  Boxes are padded and the slab test is widened by
// 4e-7 relative, so pruning is conservative.
#pragma once

#include "common.h"

namespace psdr {

constexpr int STACK_TOTAL = 64;

struct HitRec {
  int tri;
  float t, u, v;
};

// Reciprocal ray direction for the slab test.  A component of magnitude below
// 2^-64 (in particular an exact zero: axis-parallel rays, e.g. an orthographic
// camera looking along a coordinate plane) is replaced by +-2^-64, so the
// reciprocal and the products b * inv - o * inv stay finite: with an infinite
// reciprocal the fused form b * inv + (-o * inv) gives inf - inf = NaN or a
// wrongly signed infinity and rejects boxes the ray passes through.  The
// perturbation (below 2^-64 radians) is far inside the padded, widened test.
PSDR_HD f3 slab_inverse(f3 d) {
  const float m = 0x1p-64f;
  const float x = fabsf(d.x) >= m ? d.x : copysignf(m, d.x);
  const float y = fabsf(d.y) >= m ? d.y : copysignf(m, d.y);
  const float z = fabsf(d.z) >= m ? d.z : copysignf(m, d.z);
  return f3{1.0f / x, 1.0f / y, 1.0f / z};
}

// Slab test of child box `which` (0: planes a.xyz, a.w b.xy; 1: b.zw c.x,
// c.yzw): whether the ray enters it at some t in [0, tbest], with the entry
// distance in *tn_out.  tbw = tbest * 1.000001 (the caller keeps it, since
// tbest changes only with the leaf tests); the far side is widened by the same
// factor.
PSDR_HD bool slab_hit(float4 lo_hi_a, float4 lo_hi_b, float4 lo_hi_c, int which, f3 oi, f3 inv, float tbw,
                      float* tn_out) {
  float lx, ly, lz, hx, hy, hz;
  if (which == 0) {
    lx = lo_hi_a.x; ly = lo_hi_a.y; lz = lo_hi_a.z;
    hx = lo_hi_a.w; hy = lo_hi_b.x; hz = lo_hi_b.y;
  } else {
    lx = lo_hi_b.z; ly = lo_hi_b.w; lz = lo_hi_c.x;
    hx = lo_hi_c.y; hy = lo_hi_c.z; hz = lo_hi_c.w;
  }
  // t = (b - o) * inv as fma(b, inv, -o * inv); boxes are padded and the test is
  // widened, so this need not round like the exact form
  const float t0x = __builtin_fmaf(lx, inv.x, oi.x), t1x = __builtin_fmaf(hx, inv.x, oi.x);
  const float t0y = __builtin_fmaf(ly, inv.y, oi.y), t1y = __builtin_fmaf(hy, inv.y, oi.y);
  const float t0z = __builtin_fmaf(lz, inv.z, oi.z), t1z = __builtin_fmaf(hz, inv.z, oi.z);
  const float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.f));
  const float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  *tn_out = tn;
  return tn <= tf * 1.000001f && tn <= tbw;
}

// Child ids of a node.  A Karras node's children are gamma and gamma + 1 (each
// an internal node or a leaf, leaf k encoded k | 0x80000000); gamma (30 bits)
// and the two leaf flags are the low bytes of the node's first four planes
// (left lo x, y, z, left hi x), which plane_with_byte rounded outward so that
// the planes stay conservative whatever byte they carry.  The fourth 16-byte
// load a separate child-id word would cost is the saving: the traversal is
// bound as much by its vector-memory instructions as by its arithmetic.
PSDR_HD void node_children(float4 a, int* cl, int* cr) {
  const uint32_t x = __builtin_bit_cast(uint32_t, a.x), y = __builtin_bit_cast(uint32_t, a.y),
                 z = __builtin_bit_cast(uint32_t, a.z), w = __builtin_bit_cast(uint32_t, a.w);
  const int g = static_cast<int>((x & 0xffu) | ((y & 0xffu) << 8) | ((z & 0xffu) << 16) | ((w & 0x3fu) << 24));
  *cl = g | ((w & 0x40u) ? static_cast<int>(0x80000000u) : 0);
  *cr = (g + 1) | ((w & 0x80u) ? static_cast<int>(0x80000000u) : 0);
}

// node_children's fields as the traversal uses them: gamma and the two leaf
// flags (children gamma and gamma + 1).  On the device two byte permutes
// gather the four payload bytes, lo = (y.b0 << 8) | x.b0 and
// hi = (w.b0 << 24) | (z.b0 << 16), whose top two bits are the flags.
struct NodeKids {
  int g;
  bool left_leaf, right_leaf;
};
PSDR_HD NodeKids node_kids(float4 a) {
  const uint32_t x = __builtin_bit_cast(uint32_t, a.x), y = __builtin_bit_cast(uint32_t, a.y),
                 z = __builtin_bit_cast(uint32_t, a.z), w = __builtin_bit_cast(uint32_t, a.w);
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t lo = __builtin_amdgcn_perm(y, x, 0x0c0c0400u);
  const uint32_t hi = __builtin_amdgcn_perm(w, z, 0x04000c0cu);
#else
  const uint32_t lo = (x & 0xffu) | ((y & 0xffu) << 8);
  const uint32_t hi = ((z & 0xffu) << 16) | ((w & 0xffu) << 24);
#endif
  NodeKids k;
  k.g = static_cast<int>((hi & 0x3fffffffu) | lo);
  k.left_leaf = (hi & 0x40000000u) != 0u;
  k.right_leaf = static_cast<int>(hi) < 0;
  return k;
}

// A box plane p carrying one payload byte in its low mantissa bits, rounded
// outward first so that any byte keeps the plane conservative: the result is
// <= p for a low plane (hi = false) and >= p for a high one.  In the bit
// pattern of |p| (m), moving away from zero rounds m up to a multiple of 256
// (the byte then only adds); moving toward zero rounds m down to a multiple of
// 256 at least 255 below it (the byte cannot pass m); a plane within 255 ulps
// of zero becomes a tiny value of the outward sign.  The planes move by at
// most 511 ulps (3e-5 relative).  Planes are below 2^127 in magnitude (a
// triangle box padded by 1e-6 relative; larger coordinates would overflow the
// traversal's arithmetic long before).
PSDR_HD float plane_with_byte(float p, uint32_t byte, bool hi) {
  const uint32_t bits = __builtin_bit_cast(uint32_t, p);
  const uint32_t m = bits & 0x7fffffffu;
  const bool neg = (bits >> 31) != 0u && m != 0u;
  const bool away = hi ? !neg : neg;  // outward is away from zero
  uint32_t sign, mag;
  if (away) {
    sign = neg ? 0x80000000u : 0u;
    mag = (m + 255u) & ~0xffu;
  } else if (m >= 255u) {
    sign = neg ? 0x80000000u : 0u;
    mag = (m - 255u) & ~0xffu;
  } else {
    sign = hi ? 0u : 0x80000000u;
    mag = 0u;
  }
  return __builtin_bit_cast(float, sign | mag | (byte & 0xffu));
}

// The 48-byte node of child boxes bl, br ({lo xyz, hi xyz}, padded) and child
// ids cl, cr ((cr & 0x7fffffff) == (cl & 0x7fffffff) + 1, gamma < 2^30).
PSDR_HD void pack_node(const float* bl, const float* br, int cl, int cr, float4* out) {
  const uint32_t g = static_cast<uint32_t>(cl) & 0x7fffffffu;
  const uint32_t top = ((g >> 24) & 0x3fu) | (cl < 0 ? 0x40u : 0u) | (cr < 0 ? 0x80u : 0u);
  out[0] = make_float4(plane_with_byte(bl[0], g & 0xffu, false), plane_with_byte(bl[1], (g >> 8) & 0xffu, false),
                       plane_with_byte(bl[2], (g >> 16) & 0xffu, false), plane_with_byte(bl[3], top, true));
  out[1] = make_float4(bl[4], bl[5], br[0], br[1]);
  out[2] = make_float4(br[2], br[3], br[4], br[5]);
}

// Division-free rejection ahead of tri_exact (from tri_terms: det, and the
// numerators du, dv, dt of u, v, t): with s = sign(det), the
// exact test's u = du / det, v = dv / det, t = dt / det (each rounded twice)
// can only pass if s*du, s*dv >= 0 (up to values too small to round away from
// zero), s*(du + dv) <= |det|, |det| * tmin <= s*dt and s*dt <= |det| * tbest.
// The margins (2^-60 relative near zero, 2^-18 elsewhere) exceed every
// rounding error of both forms (for |det| < 2^126, i.e. coordinates far below
// 2^63), so a triangle rejected here is rejected by the exact test too; NaN
// compares false, leaving such cases to the exact test.  All six comparisons
// are evaluated (no early exit): a wave runs them for its slowest lane anyway,
// and one branch costs less than six.
PSDR_HD bool tri_maybe_hit(const TriTerms& q, float tbest) {
  const float adet = fabsf(q.det);
  const float sg = q.det < 0.f ? -1.f : 1.f;
  const float su = q.du * sg, sv = q.dv * sg, st = q.dt * sg;
  const float tiny = adet * 0x1p-60f;
  const float hi = adet * (1.0f + 0x1p-18f);
  const bool out = (su < -tiny) | (su > hi) | (sv < -tiny) | (su + sv > hi) |
                   (st < adet * (RAY_TMIN * (1.0f - 0x1p-18f))) | (st > adet * (tbest * (1.0f + 0x1p-18f)));
  return !out;
}

// Traversal statistics of a profiling launch.  NoStats compiles to nothing.
// RayStats counts the nodes fetched and the triangles tested, and keeps the
// second-smallest t among the triangles the ray hits; traversal then prunes
// against that second distance instead of the closest one, so no triangle
// that could tie with the closest hit is skipped.  The closest hit itself is
// the same either way (equal t keeps the smaller id).
struct NoStats {
  static constexpr bool on = false;
  __device__ __forceinline__ float prune(float t) const { return t; }
  __device__ __forceinline__ void node() {}
  __device__ __forceinline__ void leaf() {}
  __device__ __forceinline__ void exact() {}
  __device__ __forceinline__ void second(float) {}
};
struct RayStats {
  static constexpr bool on = true;
  float t2 = F_INF;
  int nodes = 0, leaves = 0, exacts = 0;
  __device__ __forceinline__ float prune(float) const { return t2; }
  __device__ __forceinline__ void node() { ++nodes; }
  __device__ __forceinline__ void leaf() { ++leaves; }
  __device__ __forceinline__ void exact() { ++exacts; }
  __device__ __forceinline__ void second(float t) { t2 = fminf(t2, t); }
};

// One leaf (triangle k): the prefilter, then the exact test for the lanes it
// keeps.
template <class S = NoStats>
__device__ __forceinline__ void leaf_test(const float4* __restrict__ leaf, int k, f3 o, f3 d, HitRec& h,
                                          S& st) {
  const float4 a = leaf[3 * k], b = leaf[3 * k + 1], c = leaf[3 * k + 2];
  const f3 p0{a.x, a.y, a.z}, e1{b.x, b.y, b.z}, e2{c.x, c.y, c.z};
  st.leaf();
  const TriTerms q = tri_terms(o, d, p0, e1, e2);
  if (!tri_maybe_hit(q, st.prune(h.t))) return;
  st.exact();
  float t, u, v;
  if (tri_exact(q, &t, &u, &v)) {
    const int id = __float_as_int(a.w);
    if (t < h.t || (t == h.t && id < h.tri)) {
      st.second(h.t);
      h.t = t;
      h.tri = id;
      h.u = u;
      h.v = v;
    } else if (id != h.tri) {  // a seeded traversal meets its seed's triangle again
      st.second(t);
    }
  }
}
__device__ __forceinline__ void leaf_test(const float4* __restrict__ leaf, int k, f3 o, f3 d, HitRec& h) {
  NoStats st;
  leaf_test(leaf, k, o, d, h, st);
}

// stack_lds: this block's LDS stack base ([STACK_LDS][blockDim]); spill: this
// thread's global overflow area with stride `spill_stride` between entries.
// tmax / tocc: for a visibility query (is the closest hit the target point,
// within the reference's squared-distance tolerance REAL_EPS?) hits beyond
// tmax cannot matter, and a hit before tocc already decides "not visible", so
// traversal stops there and returns that hit instead of the closest one.
// Defaults (inf, -inf) give the plain closest hit.
// Leaves [n_tris, n_all) (an environment light's bounding cube, whose
// triangles span the whole scene and would only widen every node box) are not
// in the tree: every ray tests them after the traversal.
template <class S = NoStats>
__device__ __forceinline__ void extra_leaves(const float4* __restrict__ leaf, int n_tris, int n_all, f3 o, f3 d,
                                             float tocc, HitRec& h, S& st) {
  if (h.t < tocc && h.tri >= 0) return;
  for (int k = n_tris; k < n_all; ++k) leaf_test(leaf, k, o, d, h, st);
}

// seed: a hit the ray is known to have (or null); the traversal then prunes
// against it from the start.  The result is the same as without it: the
// closest hit is the least (t, id) over all hit triangles, whatever the order
// in which they are found.
template <int STACK_LDS, class S = NoStats>
__device__ __forceinline__ HitRec closest_hit(const float4* __restrict__ node, const float4* __restrict__ leaf,
                                              int n_tris, int n_all, f3 o, f3 d, int* stack_lds, int* spill,
                                              int spill_stride, float tmax, float tocc, S& st,
                                              const HitRec* seed = nullptr) {
  HitRec h{-1, F_INF, -1.f, -1.f};
  if (!(d.x == d.x) || n_all <= 0) return h;
  h.t = tmax;
  if (seed && seed->tri >= 0) h = *seed;
  if (n_tris <= 1) {
    if (n_tris == 1) leaf_test(leaf, 0, o, d, h, st);
    extra_leaves(leaf, n_tris, n_all, o, d, tocc, h, st);
    if (h.tri < 0) h.t = F_INF;
    return h;
  }
  const f3 inv = slab_inverse(d);
  const f3 oi = f3{-(o.x * inv.x), -(o.y * inv.y), -(o.z * inv.z)};
  const int lane = threadIdx.x, nthr = blockDim.x;
  int sp = 0;
  int ni = 0;
  const int n_internal = n_tris - 1;
  float tbw = st.prune(h.t) * 1.000001f;  // changes only with the leaf tests
  for (;;) {
    if (static_cast<unsigned>(ni) >= static_cast<unsigned>(n_internal)) break;  // malformed tree guard
    const float4 a = node[3 * ni], b = node[3 * ni + 1], c = node[3 * ni + 2];
    st.node();
    const NodeKids kid = node_kids(a);
    const int cl = kid.g, cr = kid.g + 1;
    float tl, tr;
    const bool el = slab_hit(a, b, c, 0, oi, inv, tbw, &tl);
    const bool er = slab_hit(a, b, c, 1, oi, inv, tbw, &tr);
    // leaf children the ray enters: a lane's first one (left or right) in one
    // pass, so lanes entering only a left leaf and lanes entering only a right
    // one test together; the right one of lanes entering both in a second pass.
    // The closest hit is the least (t, id) however the leaves are ordered.
    const bool ll = el && kid.left_leaf, lr = er && kid.right_leaf;
    if (ll || lr) {
      const int k0 = ll ? cl : cr;
      if (k0 < n_tris) leaf_test(leaf, k0, o, d, h, st);
      if (ll && lr) {
        if (cr < n_tris) leaf_test(leaf, cr, o, d, h, st);
      }
      tbw = st.prune(h.t) * 1.000001f;
    }
    if (h.t < tocc && h.tri >= 0) break;  // occluded before the target
    const bool hl = el && !kid.left_leaf, hr = er && !kid.right_leaf;
    if (hl && hr) {
      int nearc = cl, farc = cr;
      if (tr < tl) {
        nearc = cr;
        farc = cl;
      }
      if (sp < STACK_LDS) stack_lds[sp * nthr + lane] = farc;
      else if (sp < STACK_TOTAL) spill[(sp - STACK_LDS) * spill_stride] = farc;
      if (sp < STACK_TOTAL) ++sp;
      ni = nearc;
    } else if (hl) {
      ni = cl;
    } else if (hr) {
      ni = cr;
    } else {
      if (sp == 0) break;
      --sp;
      // the spill read is volatile so that the compiler keeps two loads: merged,
      // they become one flat (generic address) load on every pop, which is
      // slower than the LDS read taken by all but the deepest entries
      if (sp < STACK_LDS) ni = stack_lds[sp * nthr + lane];
      else ni = *reinterpret_cast<volatile const int*>(spill + (sp - STACK_LDS) * spill_stride);
    }
  }
  extra_leaves(leaf, n_tris, n_all, o, d, tocc, h, st);
  if (h.tri < 0) h.t = F_INF;
  return h;
}

template <int STACK_LDS>
__device__ __forceinline__ HitRec closest_hit(const float4* __restrict__ node, const float4* __restrict__ leaf,
                                              int n_tris, int n_all, f3 o, f3 d, int* stack_lds, int* spill,
                                              int spill_stride, float tmax = F_INF, float tocc = -F_INF) {
  NoStats st;
  return closest_hit<STACK_LDS>(node, leaf, n_tris, n_all, o, d, stack_lds, spill, spill_stride, tmax, tocc, st);
}

// A BSDF-sampled ray only needs its closest hit when that hit is an emitting
// triangle (src/path_integrator.cc:134-160 adds light only then).  For such a
// ray (seek) the scene's few emitting triangles (seek leaves) are tested first:
// a ray that hits none of them cannot have an emitting closest hit, so *skip is
// set and the returned miss (tri -1) stands for its hit; otherwise the nearest
// emitting hit is returned as the seed of the ordinary closest-hit traversal.
// Without seek the result is an empty seed.
template <class S>
__device__ __forceinline__ HitRec seek_lights(const float4* __restrict__ seek_leaf, int n_seek, bool seek, f3 o,
                                              f3 d, bool* skip, S& st) {
  HitRec h{-1, F_INF, -1.f, -1.f};
  *skip = false;
  if (!seek || !(d.x == d.x)) return h;
  for (int m = 0; m < n_seek; ++m) leaf_test(seek_leaf, m, o, d, h, st);
  *skip = h.tri < 0;
  return h;
}

// A closest hit within 4 ulp of the next triangle the ray hits (an order
// between them that another triangle test could flip), and one whose
// barycentrics lie within 1e-6 of an edge (a neighbour could claim it).
__device__ __forceinline__ bool near_tie(const HitRec& h, float t2) {
  return h.tri >= 0 && t2 < F_INF && __float_as_int(t2) - __float_as_int(h.t) <= 4;
}
__device__ __forceinline__ bool grazing(const HitRec& h) {
  return h.tri >= 0 && fminf(fminf(h.u, h.v), 1.0f - h.u - h.v) < 1e-6f;
}

// Window of a visibility query toward a target at distance t: a hit at least
// 4e-3 (plus 1e-5 relative, for the direction's rounding) nearer or farther
// than the target lies farther from it than sqrt(REAL_EPS) = 3.2e-3.
__device__ __forceinline__ float vis_tmax(float t) { return t + (4e-3f + 1e-5f * t); }
__device__ __forceinline__ float vis_tocc(float t) { return t - (4e-3f + 1e-5f * t); }

}  // namespace psdr
