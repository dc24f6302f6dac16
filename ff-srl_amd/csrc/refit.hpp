// Bottom-up BVH refit of one environment (refitBVHLevel, bvh.py:30-104).
//
// Nodes of a mesh are pre-grouped for the refit: all leaves first (they only
// read positions), then the inner nodes level by level, deepest first, each
// group one contiguous range with the node's (left child or first triangle,
// count) beside its id.  The refit walks the groups with a workgroup barrier
// between them instead of launching every node once per level as
// refitBVHLoop does (bvh.py:419-440).  Leaf bounds start at +-1e6
// and take min/max over the three vertices of each of the leaf's triangles in
// the reference's operand order; inner nodes take the union of children
// (left, left + 1).  min/max select exact inputs, so the bounds are bit-exact.
#pragma once

#include "common.hpp"

namespace ffsrl {

// Refit of entry k of the env's mesh's refit groups.
__device__ __forceinline__ void refit_node(const ffsrl_batch& b, const ffsrl_env_desc& e, int k) {
  const int64_t kb = (int64_t)e.nodebase + k;
  const int node = b.level_nodes[kb];
  const int2 info = ldi2(b.level_info, kb);
  const int64_t vb = e.vbase;
  const int n = info.y;
  float4 lo, hi;
  if (n > 0) {
    const int first = info.x;
    f3 mn = mk(1e6f, 1e6f, 1e6f), mx = mk(-1e6f, -1e6f, -1e6f);
    for (int k = 0; k < n; ++k) {
      const int4 t = ldi4(b.tris, (int64_t)e.fbase + first + k);
      const f3 A = xyz(ld4(b.x, vb + t.x)), B = xyz(ld4(b.x, vb + t.y)), C = xyz(ld4(b.x, vb + t.z));
      mn = vmin(A, vmin(B, vmin(C, mn)));
      mx = vmax(A, vmax(B, vmax(C, mx)));
    }
    lo = make_float4(mn.x, mn.y, mn.z, __int_as_float(first));
    hi = make_float4(mx.x, mx.y, mx.z, __int_as_float(n));
  } else {
    const int l = info.x;
    const int64_t sb = e.boundbase;
    const float4 al = ld4(b.bvh_lo, sb + l), ar = ld4(b.bvh_lo, sb + l + 1);
    const float4 bl = ld4(b.bvh_hi, sb + l), br = ld4(b.bvh_hi, sb + l + 1);
    const f3 mn = vmin(xyz(al), xyz(ar)), mx = vmax(xyz(bl), xyz(br));
    lo = make_float4(mn.x, mn.y, mn.z, __int_as_float(l));
    hi = make_float4(mx.x, mx.y, mx.z, __int_as_float(0));
  }
  st4(b.bvh_lo, (int64_t)e.boundbase + node, lo);
  st4(b.bvh_hi, (int64_t)e.boundbase + node, hi);
}

template <int BLOCK>
__device__ void refit_env(const ffsrl_batch& b, const ffsrl_env_desc& e) {
  for (int L = 0; L < e.nlvl; ++L) {
    const int beg = b.level_off[e.lvlbase + L], end = b.level_off[e.lvlbase + L + 1];
    for (int k = beg + (int)threadIdx.x; k < end; k += BLOCK) refit_node(b, e, k);
    __syncthreads();
  }
}

}  // namespace ffsrl
