// The renderer's ray / node box test: intersectAABB (render.py:588-601) and
// the traversal's decision "hit && tmin < tMin" (render.py:447-452), with a
// fast form that gives the same decision.  Host and device code, so that the
// fast form can be checked against the exact one on the host.
#pragma once

#include "common.hpp"

namespace ffsrl {

// intersectAABB (render.py:588-601) expression by expression: slab values
// (lo - o) / d, then min / max by compare-select.
__host__ __device__ inline int isect_aabb(f3 o, f3 d, f3 lo, f3 hi, float* tmin_out) {
  const float t1 = (lo.x - o.x) / d.x, t2 = (hi.x - o.x) / d.x;
  const float t3 = (lo.y - o.y) / d.y, t4 = (hi.y - o.y) / d.y;
  const float t5 = (lo.z - o.z) / d.z, t6 = (hi.z - o.z) / d.z;
  const float tmin = fmax_(fmax_(fmin_(t1, t2), fmin_(t3, t4)), fmin_(t5, t6));
  const float tmax = fmin_(fmin_(fmax_(t1, t2), fmax_(t3, t4)), fmax_(t5, t6));
  *tmin_out = tmin;
  return (tmax >= 0.0f) && (tmax >= tmin);
}

__host__ __device__ inline bool exact_visit(f3 o, f3 d, f3 lo, f3 hi, float tMin) {
  float tb;
  return isect_aabb(o, d, lo, hi, &tb) && tb < tMin;
}

__host__ __device__ inline bool close_rel(float a, float b) { return fabsf(a - b) <= 1e-6f * (fabsf(a) + fabsf(b)); }

// IEEE 754-2019 minimum / maximum (v_minimum / v_maximum): NaN propagates.
__host__ __device__ inline float vmin_(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__host__ __device__ inline float vmax_(float a, float b) { return __builtin_elementwise_maximum(a, b); }

// A direction component in (0, 1e-30) in magnitude sends every test of the
// ray to the exact form (then |1/d| <= 1e30 in the fast form).
__host__ __device__ inline bool ray_dir_ok(f3 d) {
  auto ok = [](float c) { return (c == 0.0f) | (fabsf(c) >= 1e-30f); };
  return ok(d.x) & ok(d.y) & ok(d.z);
}

// Fast form (products): slab values (lo - o) * (1/d), within 3 ulp of the
// quotients; special values (d = 0) give the same infinities and NaNs.  When
// a comparison of the decision is closer than 1e-6 relative, or a slab value
// is not finite, the decision is taken from the exact form.
//
// min/max are IEEE minimum/maximum, which propagate NaN: a NaN among the six
// products makes tmin and tmax NaN, and |NaN| > 1e-30 is false, so such a
// node goes to the exact form.  Otherwise they select exactly what the
// reference's compare-select min/max select, up to the sign of a zero, which
// no comparison below can see.  A product that overflows where the quotient
// does not could change a selection only among values beyond 1e37: such tmin
// / tmax go to the exact form, and so do rays with a direction component in
// (0, 1e-30) (ray_ok).  Other infinities come from a direction component that
// is exactly zero, where the quotient is the same infinity.
__host__ __device__ inline bool visit_node(f3 o, f3 d, f3 inv, f3 lo, f3 hi, float tMin, bool ray_ok) {
  const float t1 = (lo.x - o.x) * inv.x, t2 = (hi.x - o.x) * inv.x;
  const float t3 = (lo.y - o.y) * inv.y, t4 = (hi.y - o.y) * inv.y;
  const float t5 = (lo.z - o.z) * inv.z, t6 = (hi.z - o.z) * inv.z;
  const float tmin = vmax_(vmax_(vmin_(t1, t2), vmin_(t3, t4)), vmin_(t5, t6));
  const float tmax = vmin_(vmin_(vmax_(t1, t2), vmax_(t3, t4)), vmax_(t5, t6));
  // every condition is evaluated (bitwise &, no short circuit), so the whole
  // fast form is one basic block with a single branch
  const bool robust = ray_ok & (fabsf(tmax) > 1e-30f) & (fabsf(tmax) < 1e37f) & (fabsf(tmin) < 1e37f) &
                      !close_rel(tmax, tmin) & !close_rel(tmin, tMin);
  if (robust) return (tmax >= 0.0f) & (tmax >= tmin) & (tmin < tMin);
  return exact_visit(o, d, lo, hi, tMin);
}

}  // namespace ffsrl
