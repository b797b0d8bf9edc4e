// Host-only check (built with hipcc, runs on the CPU): the traversal's slab
// test (slab_inverse + slab_hit in traverse.h) never rejects a node box whose
// content the ray meets at some t in [0, tbest].  Node boxes are the union of
// triangle boxes padded by (|lo| + |hi|) * 1e-6 + 1e-30 per axis (scene.hip), so
// the content lies in the box shrunk by that pad; the reference is an exact
// test in double on the shrunk box, a zero direction component meaning "inside
// the slab iff the origin is".  Rays with exactly zero (and signed-zero, and
// denormal) direction components, origins on the content's faces, boxes over
// six orders of magnitude.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../torch_psdr_amd/csrc/traverse.h"

using namespace psdr;

static bool exact_hit(const double lo[3], const double hi[3], const double o[3], const double d[3], double tbest) {
  double tn = 0.0, tf = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0.0) {
      if (o[a] < lo[a] || o[a] > hi[a]) return false;
      continue;
    }
    double t0 = (lo[a] - o[a]) / d[a], t1 = (hi[a] - o[a]) / d[a];
    if (t0 > t1) std::swap(t0, t1);
    tn = std::max(tn, t0);
    tf = std::min(tf, t1);
  }
  return tn <= tf && tn <= tbest;
}

int main(int argc, char** argv) {
  const long long iters = argc > 1 ? std::atoll(argv[1]) : 2000000;
  std::mt19937_64 g(7);
  std::uniform_real_distribution<float> U(-1.f, 1.f);
  long long met = 0, bad = 0, pruned = 0;
  for (long long it = 0; it < iters; ++it) {
    const float sc = std::pow(10.f, U(g) * 3.f);
    float lo[3], hi[3], o[3], d[3];
    float clo[3], chi[3];  // content box; lo/hi = padded node box
    for (int a = 0; a < 3; ++a) {
      const float c = U(g) * sc, h = std::fabs(U(g)) * sc * 0.5f;
      clo[a] = c - h;
      chi[a] = c + h;
      const float pad = (std::fabs(clo[a]) + std::fabs(chi[a])) * 1e-6f + 1e-30f;
      lo[a] = clo[a] - pad;
      hi[a] = chi[a] + pad;
      o[a] = U(g) * sc * 2.f;
      d[a] = U(g);
    }
    // zero, negative-zero and denormal components; origins on faces
    const int z = static_cast<int>(it % 7);
    if (z < 3) d[z] = (it & 8) ? -0.0f : 0.0f;
    if (z == 3) {
      d[0] = 0.f;
      d[1] = -0.f;
    }
    if (z == 4) d[static_cast<int>(it / 7) % 3] = 1e-42f * ((it & 16) ? -1.f : 1.f);
    if (it % 5 == 0) o[static_cast<int>(it / 5) % 3] = (it & 32) ? clo[static_cast<int>(it / 5) % 3]
                                                                  : chi[static_cast<int>(it / 5) % 3];
    if (d[0] == 0.f && d[1] == 0.f && d[2] == 0.f) d[2] = 1.f;
    const float tbest = (it % 3 == 0) ? INFINITY : std::fabs(U(g)) * sc * 4.f;
    const f3 df{d[0], d[1], d[2]}, of{o[0], o[1], o[2]};
    const f3 inv = slab_inverse(df);
    const f3 oi = f3{-(of.x * inv.x), -(of.y * inv.y), -(of.z * inv.z)};
    const float4 a4 = make_float4(lo[0], lo[1], lo[2], hi[0]);
    const float4 b4 = make_float4(hi[1], hi[2], 0.f, 0.f);
    const float4 c4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float tn_e;
    const bool te = slab_hit(a4, b4, c4, 0, oi, inv, tbest * 1.000001f, &tn_e);
    double lod[3], hid[3], od[3], dd[3];
    for (int a = 0; a < 3; ++a) {
      lod[a] = clo[a];
      hid[a] = chi[a];
      od[a] = o[a];
      dd[a] = d[a];
    }
    const bool ex = exact_hit(lod, hid, od, dd, tbest);
    if (ex) {
      ++met;
      if (!te) {
        ++bad;
        if (bad <= 5)
          std::printf("rejected: o %g %g %g d %g %g %g lo %g %g %g hi %g %g %g tbest %g\n", o[0], o[1], o[2], d[0],
                      d[1], d[2], lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], tbest);
      }
    } else if (!te) {
      ++pruned;
    }
  }
  std::printf("boxes_met %lld false_rejections %lld pruned %lld of %lld\n", met, bad, pruned, iters);
  return bad == 0 && pruned > iters / 10 ? 0 : 1;
}
