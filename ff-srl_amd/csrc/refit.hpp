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

// Refit of one node given its refit-group entry (node id; left child or
// first triangle, count).  COH: positions are read with ld4c (written by
// other workgroups during the same kernel).
template <bool COH = false>
__device__ __forceinline__ void refit_node_info(const ffsrl_batch& b, const ffsrl_env_desc& e, int node, int2 info) {
  const int64_t vb = e.vbase;
  const int n = info.y;
  float4 lo, hi;
  if (n > 0) {
    const int first = info.x;
    f3 mn = mk(1e6f, 1e6f, 1e6f), mx = mk(-1e6f, -1e6f, -1e6f);
    for (int k = 0; k < n; ++k) {
      const int4 t = ldi4(b.tris, (int64_t)e.fbase + first + k);
      const f3 A = xyz(ldp<COH>(b.x, vb + t.x)), B = xyz(ldp<COH>(b.x, vb + t.y)), C = xyz(ldp<COH>(b.x, vb + t.z));
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

// Refit of entry k of the env's mesh's refit groups.
template <bool COH = false>
__device__ __forceinline__ void refit_node(const ffsrl_batch& b, const ffsrl_env_desc& e, int k) {
  const int64_t kb = (int64_t)e.nodebase + k;
  refit_node_info<COH>(b, e, b.level_nodes[kb], ldi2(b.level_info, kb));
}

// U entries of one group (k, k + stride, ...): every load of the U entries is
// issued before the first store, so their latencies overlap.  A group holds
// only leaves or only inner nodes; single-triangle leaves and inner nodes take
// this path, anything else goes entry by entry through refit_node.
template <int U, bool COH = false>
__device__ __forceinline__ void refit_entries(const ffsrl_batch& b, const ffsrl_env_desc& e, int k, int stride) {
  int node[U];
  int2 info[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t kb = (int64_t)e.nodebase + k + u * stride;
    node[u] = b.level_nodes[kb];
    info[u] = ldi2(b.level_info, kb);
  }
  bool leaf1 = true, inner = true;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    leaf1 = leaf1 && info[u].y == 1;
    inner = inner && info[u].y == 0;
  }
  const int64_t vb = e.vbase, sb = e.boundbase;
  float4 lo[U], hi[U];
  if (leaf1) {
    int4 t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = ldi4(b.tris, (int64_t)e.fbase + info[u].x);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const f3 A = xyz(ldp<COH>(b.x, vb + t[u].x)), B = xyz(ldp<COH>(b.x, vb + t[u].y)), C = xyz(ldp<COH>(b.x, vb + t[u].z));
      const f3 mn = vmin(A, vmin(B, vmin(C, mk(1e6f, 1e6f, 1e6f))));
      const f3 mx = vmax(A, vmax(B, vmax(C, mk(-1e6f, -1e6f, -1e6f))));
      lo[u] = make_float4(mn.x, mn.y, mn.z, __int_as_float(info[u].x));
      hi[u] = make_float4(mx.x, mx.y, mx.z, __int_as_float(1));
    }
  } else if (inner) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int l = info[u].x;
      const float4 al = ld4(b.bvh_lo, sb + l), ar = ld4(b.bvh_lo, sb + l + 1);
      const float4 bl = ld4(b.bvh_hi, sb + l), br = ld4(b.bvh_hi, sb + l + 1);
      const f3 mn = vmin(xyz(al), xyz(ar)), mx = vmax(xyz(bl), xyz(br));
      lo[u] = make_float4(mn.x, mn.y, mn.z, __int_as_float(l));
      hi[u] = make_float4(mx.x, mx.y, mx.z, __int_as_float(0));
    }
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) refit_node<COH>(b, e, k + u * stride);
    return;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    st4(b.bvh_lo, sb + node[u], lo[u]);
    st4(b.bvh_hi, sb + node[u], hi[u]);
  }
}

constexpr int kMaxRefitGroups = 64;  // leaf group + inner levels (BVH depth <= 32)

// The groups in order with a barrier after each.  The group offsets are read
// once into LDS, and each thread loads its first entry of the next group
// before the barrier (entries do not depend on bounds), so a group costs the
// round trip of the child bounds and the barrier.
template <int BLOCK, bool COH = false>
__device__ void refit_env(const ffsrl_batch& b, const ffsrl_env_desc& e) {
  __shared__ int off[kMaxRefitGroups + 1];
  const int ng = e.nlvl;
  if (ng > kMaxRefitGroups) {  // deeper trees than the renderer accepts: plain group walk
    for (int L = 0; L < ng; ++L) {
      const int beg = b.level_off[e.lvlbase + L], end = b.level_off[e.lvlbase + L + 1];
      for (int k = beg + (int)threadIdx.x; k < end; k += BLOCK) refit_node<COH>(b, e, k);
      __syncthreads();
    }
    return;
  }
  for (int i = threadIdx.x; i <= ng; i += BLOCK) off[i] = b.level_off[e.lvlbase + i];
  __syncthreads();
  const int tid = threadIdx.x;
  int node = 0;
  int2 info = make_int2(0, 0);
  if (ng > 0 && off[0] + tid < off[1]) {
    const int64_t kb = (int64_t)e.nodebase + off[0] + tid;
    node = b.level_nodes[kb];
    info = ldi2(b.level_info, kb);
  }
  for (int L = 0; L < ng; ++L) {
    const int beg = off[L], end = off[L + 1];
    int k = beg + tid;
    if (k < end) {
      refit_node_info<COH>(b, e, node, info);
      k += BLOCK;
    }
    for (; k + 3 * BLOCK < end; k += 4 * BLOCK) refit_entries<4, COH>(b, e, k, BLOCK);
    if (k + BLOCK < end) {
      refit_entries<2, COH>(b, e, k, BLOCK);
      k += 2 * BLOCK;
    }
    if (k < end) refit_node<COH>(b, e, k);
    if (L + 1 < ng && off[L + 1] + tid < off[L + 2]) {  // the next group's first entry
      const int64_t kb = (int64_t)e.nodebase + off[L + 1] + tid;
      node = b.level_nodes[kb];
      info = ldi2(b.level_info, kb);
    }
    __syncthreads();
  }
}

}  // namespace ffsrl
