// Host-only check (built with hipcc, runs on the CPU) of the 48-byte LBVH node
// (traverse.h pack_node / node_children / node_kids / plane_with_byte / slab_hit):
//  * plane_with_byte never moves a plane inward, whatever byte it carries: a low
//    plane comes out <= p, a high one >= p (normal, tiny, denormal, signed-zero
//    planes of magnitude below 2^127), and never becomes inf or NaN;
//  * a packed node's child ids decode to the ids it was packed from (gamma up to
//    2^30 - 1, every combination of leaf flags), its untouched planes are the
//    input planes bit for bit, and the slab test on the packed node enters every
//    box that the test on the exact planes enters.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../torch_psdr_amd/csrc/traverse.h"

using namespace psdr;

static float bits_to_float(uint32_t b) {
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

int main(int argc, char** argv) {
  const long long iters = argc > 1 ? std::atoll(argv[1]) : 2000000;
  std::mt19937_64 g(11);
  long long bad_plane = 0, bad_ids = 0, bad_copy = 0, bad_slab = 0, entered = 0;
  for (long long it = 0; it < iters; ++it) {
    // planes: random bit patterns (finite), values near zero, signed zeros
    float p;
    const int kind = static_cast<int>(it % 5);
    if (kind == 0) {
      uint32_t b = static_cast<uint32_t>(g());
      if (((b >> 23) & 0xffu) >= 0xfeu) b &= ~(1u << 30);  // |p| < 2^127 (no inf / NaN)
      p = bits_to_float(b);
    } else if (kind == 1) {
      p = bits_to_float(static_cast<uint32_t>(g() % 600u) | ((g() & 1u) ? 0x80000000u : 0u));
    } else if (kind == 2) {
      p = (g() & 1u) ? -0.0f : 0.0f;
    } else {
      std::uniform_real_distribution<float> U(-1.f, 1.f);
      p = U(g) * std::pow(10.f, U(g) * 6.f);
    }
    const uint32_t byte = static_cast<uint32_t>(g() & 0xffu);
    const bool hi = (g() & 1u) != 0u;
    const float q = plane_with_byte(p, byte, hi);
    if (!std::isfinite(q) || (hi ? !(q >= p) : !(q <= p)) || (__builtin_bit_cast(uint32_t, q) & 0xffu) != byte)
      ++bad_plane;

    // nodes
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    const float sc = std::pow(10.f, U(g) * 3.f);
    float bl[6], br[6];
    for (int a = 0; a < 3; ++a) {
      const float c0 = U(g) * sc, h0 = std::fabs(U(g)) * sc;
      const float c1 = U(g) * sc, h1 = std::fabs(U(g)) * sc;
      bl[a] = c0 - h0;
      bl[3 + a] = c0 + h0;
      br[a] = c1 - h1;
      br[3 + a] = c1 + h1;
    }
    const int gamma = static_cast<int>(g() % (1u << 30) - (g() & 1u));
    const int gm = gamma < 0 ? 0 : (gamma > (1 << 30) - 2 ? (1 << 30) - 2 : gamma);
    const int cl = gm | ((g() & 1u) ? static_cast<int>(0x80000000u) : 0);
    const int cr = (gm + 1) | ((g() & 1u) ? static_cast<int>(0x80000000u) : 0);
    float4 nd[3];
    pack_node(bl, br, cl, cr, nd);
    int dl, dr;
    node_children(nd[0], &dl, &dr);
    if (dl != cl || dr != cr) ++bad_ids;
    const NodeKids kk = node_kids(nd[0]);
    if (kk.g != (cl & 0x7fffffff) || kk.left_leaf != (cl < 0) || kk.right_leaf != (cr < 0)) ++bad_ids;
    const float tail[8] = {nd[1].x, nd[1].y, nd[1].z, nd[1].w, nd[2].x, nd[2].y, nd[2].z, nd[2].w};
    const float want[8] = {bl[4], bl[5], br[0], br[1], br[2], br[3], br[4], br[5]};
    if (std::memcmp(tail, want, sizeof tail) != 0) ++bad_copy;
    // slab test: exact planes vs packed planes, same ray
    const f3 o{U(g) * sc * 2.f, U(g) * sc * 2.f, U(g) * sc * 2.f};
    f3 d{U(g), U(g), U(g)};
    if (d.x == 0.f && d.y == 0.f && d.z == 0.f) d.z = 1.f;
    const f3 inv = slab_inverse(d);
    const f3 oi = f3{-(o.x * inv.x), -(o.y * inv.y), -(o.z * inv.z)};
    const float tbest = (it % 3 == 0) ? INFINITY : std::fabs(U(g)) * sc * 4.f;
    const float4 ea = make_float4(bl[0], bl[1], bl[2], bl[3]);
    const float4 eb = make_float4(bl[4], bl[5], br[0], br[1]);
    const float4 ec = make_float4(br[2], br[3], br[4], br[5]);
    for (int w = 0; w < 2; ++w) {
      float tn_e, tn_p;
      const bool te = slab_hit(ea, eb, ec, w, oi, inv, tbest * 1.000001f, &tn_e);
      const bool tp = slab_hit(nd[0], nd[1], nd[2], w, oi, inv, tbest * 1.000001f, &tn_p);
      if (te) {
        ++entered;
        if (!tp) ++bad_slab;
      }
    }
  }
  std::printf("iters %lld entered %lld bad_plane %lld bad_ids %lld bad_copy %lld bad_slab %lld\n", iters, entered,
              bad_plane, bad_ids, bad_copy, bad_slab);
  const bool ok = bad_plane == 0 && bad_ids == 0 && bad_copy == 0 && bad_slab == 0 && entered > iters / 10;
  std::printf(ok ? "node_fuzz ok\n" : "node_fuzz FAILED\n");
  return ok ? 0 : 1;
}
