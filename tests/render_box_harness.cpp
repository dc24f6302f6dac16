// Host check of the renderer's fast ray / box decision (render_box.hpp:
// visit_node) against the exact form (exact_visit, the
// reference's intersectAABB with "tmin < tMin"), on seeded random and
// adversarial cases: rays with zero or tiny direction components, boxes whose
// faces pass through points of the ray (grazing hits, tmin == tmax), tMin a
// few ulps either side of the exact entry distance, box faces at the ray
// origin (cancellation in lo - o), far origins, tiny and inverted boxes.
// Prints the case count and the mismatches; exit status 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../ff-srl_amd/csrc/render_box.hpp"

using namespace ffsrl;

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  const unsigned long long seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1ull;
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  auto uf = [&](double a, double b) { return (float)(a + (b - a) * U(g)); };
  auto ui = [&](int k) { return (int)(U(g) * k); };
  long mism = 0, hits = 0;
  for (long i = 0; i < n; ++i) {
    const int kind = (int)(i % 10);
    const float so = powf(10.0f, uf(-3.0, kind == 6 ? 6.0 : 2.0));
    f3 o = mk(uf(-1, 1) * so, uf(-1, 1) * so, uf(-1, 1) * so);
    if (ui(8) == 0) o.y = 0.0f;
    f3 d = normalize(mk(uf(-1, 1), uf(-1, 1), uf(-1, 1)));
    if (kind == 1) {
      float* c = ui(3) == 0 ? &d.x : ui(2) ? &d.y : &d.z;
      *c = ui(3) == 0 ? 0.0f : (ui(2) ? 1.0f : -1.0f) * powf(10.0f, -uf(3, 36));
    }
    // a point of the ray (in float) and a box around it
    const float t = (ui(10) == 0 ? -1.0f : 1.0f) * powf(10.0f, uf(-3, 3));
    const f3 p = mk(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    const float sz = fabsf(t) * powf(10.0f, kind == 5 ? uf(-8, -5) : uf(-4, 0.5));
    f3 lo = mk(p.x - sz * uf(0, 1), p.y - sz * uf(0, 1), p.z - sz * uf(0, 1));
    f3 hi = mk(p.x + sz * uf(0, 1), p.y + sz * uf(0, 1), p.z + sz * uf(0, 1));
    if (kind == 2) {
      // faces through ray points: the ray enters / leaves on an edge or corner
      const float t2 = t * (1.0f + uf(-1e-3, 1e-3) * (ui(2) ? 1.0f : 0.0f));
      const f3 q = mk(o.x + t2 * d.x, o.y + t2 * d.y, o.z + t2 * d.z);
      const int a = ui(3), b2 = (a + 1 + ui(2)) % 3;
      float* L = &lo.x;
      float* H = &hi.x;
      const float* Q = &q.x;
      if (ui(2)) L[a] = Q[a]; else H[a] = Q[a];
      if (ui(2)) L[b2] = Q[b2]; else H[b2] = Q[b2];
    }
    if (kind == 4) {
      // a face at the ray origin's coordinate
      float* L = &lo.x;
      const float* O = &o.x;
      const int a = ui(3);
      L[a] = O[a];
    }
    if (kind == 7) {
      // anywhere, possibly inverted
      const float s = so * 3.0f;
      lo = mk(uf(-1, 1) * s, uf(-1, 1) * s, uf(-1, 1) * s);
      hi = mk(lo.x + uf(-0.2, 1) * s, lo.y + uf(-0.2, 1) * s, lo.z + uf(-0.2, 1) * s);
    }
    float tMin = 1e6f;
    if (kind == 3 || kind == 8 || kind == 9) {
      // tMin within a few ulps of the exact entry distance
      float tb;
      (void)isect_aabb(o, d, lo, hi, &tb);
      if (std::isfinite(tb)) {
        tMin = tb;
        const int k = ui(5);
        const float dir = ui(2) ? INFINITY : -INFINITY;
        for (int j = 0; j < k; ++j) tMin = nextafterf(tMin, dir);
      } else {
        tMin = uf(0, 10);
      }
    } else if (ui(3) == 0) {
      tMin = fabsf(t) * uf(0.5, 2.0);
    }
    const bool ex = exact_visit(o, d, lo, hi, tMin);
    const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const bool fast = visit_node(o, d, inv, lo, hi, tMin, ray_dir_ok(d));
    hits += ex;
    if (fast != ex) {
      ++mism;
      if (mism <= 5)
        printf("mismatch kind %d: o (%.9g %.9g %.9g) d (%.9g %.9g %.9g) lo (%.9g %.9g %.9g) hi (%.9g %.9g %.9g) "
               "tMin %.9g exact %d\n",
               kind, o.x, o.y, o.z, d.x, d.y, d.z, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, tMin, (int)ex);
    }
  }
  printf("cases %ld hits %ld mismatches %ld\n", n, hits, mism);
  return mism ? 1 : 0;
}
