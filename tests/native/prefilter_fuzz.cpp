// Host-only check (built with hipcc, runs on the CPU): the division-free
// triangle rejection in traverse.h never rejects a triangle that the exact
// Moller-Trumbore test accepts with t <= tbest.  Random triangles over six
// orders of magnitude, near-degenerate ones, rays aimed exactly at edges and
// vertices and rays starting at t = tmin.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../torch_psdr_amd/csrc/traverse.h"

using namespace psdr;

int main(int argc, char** argv) {
  const long long iters = argc > 1 ? std::atoll(argv[1]) : 2000000;
  std::mt19937_64 g(1);
  std::uniform_real_distribution<float> U(-1.f, 1.f);
  long long hits = 0, bad = 0, rejected = 0;
  for (long long it = 0; it < iters; ++it) {
    const float sc = std::pow(10.f, U(g) * 3.f);
    const f3 p0{U(g) * sc, U(g) * sc, U(g) * sc};
    const f3 e1{U(g) * sc, U(g) * sc, U(g) * sc};
    f3 e2{U(g) * sc, U(g) * sc, U(g) * sc};
    if (it % 7 == 0) e2 = e1 * (1.f + U(g) * 1e-6f);
    f3 o{U(g) * 3, U(g) * 3, U(g) * 3};
    float a = U(g) * 0.6f + 0.3f, b = U(g) * 0.6f + 0.3f;
    if (it % 5 == 0) a = 0.f;
    if (it % 11 == 0) b = 1.f - a;
    if (it % 13 == 0) {
      a = 1.f;
      b = 0.f;
    }
    const f3 target = p0 + e1 * a + e2 * b;
    const f3 d = normalize(target - o);
    if (it % 17 == 0) o = target - d * RAY_TMIN;
    const float tbest = (it % 3 == 0) ? INFINITY : std::fabs(U(g)) * 5.f;
    float t, u, v;
    const TriTerms q = tri_terms(o, d, p0, e1, e2);
    const bool hit = tri_exact(q, &t, &u, &v) && t <= tbest;
    const bool maybe = tri_maybe_hit(q, tbest);
    hits += hit;
    rejected += !maybe;
    if (hit && !maybe) {
      if (bad < 5) std::printf("false rejection at %lld: t %g u %g v %g tbest %g\n", it, t, u, v, tbest);
      ++bad;
    }
  }
  std::printf("iterations %lld hits %lld rejected %lld false_rejections %lld\n", iters, hits, rejected, bad);
  return bad != 0;
}
