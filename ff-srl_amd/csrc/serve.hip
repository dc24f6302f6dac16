// Request kernels of the serving loop: apply a tick's reset / step requests to
// the listed environments, and reduce each listed environment to its
// observation scalars.  Both run one workgroup per listed environment, so a
// tick of mixed requests is two launches (plus the frame and the render).
//
// Reset follows resetFixedModel (modelDO.py:1589-1622): positions restored,
// velocities and grasp weights zeroed, instrument points restored; the env's
// BVH is refit in the same workgroup so the reset observation renders the
// restored mesh.  A step moves the instrument points by the request's delta
// (clamped to the workspace, as applyCartesianActionsInWorkspace clamps,
// modelDO.py:1385-1404) and, when the request opens or closes the clamp,
// reselects the grasped vertices like updateClampsDragKernel
// (modelDO.py:247-279): drag = 1 for vertices within `radius` of the drag
// point tip[env*4+3] when closed, 0 when open.
#include "common.hpp"
#include "refit.hpp"

namespace ffsrl {

constexpr int kServeBlock = 256;

__global__ __launch_bounds__(kServeBlock) void k_apply_requests(ffsrl_batch b, ffsrl_requests r, float radius) {
  const int i = blockIdx.x;
  const int env = r.envs[i];
  const ffsrl_env_desc e = b.envs[env];
  const int64_t vb = e.vbase;
  const int tid = threadIdx.x;
  __shared__ float tip3[3];
  if (r.kind[i] == FFSRL_REQ_RESET) {
    for (int v = tid; v < e.nv; v += kServeBlock) {
      const float4 x0 = ld4(r.x0, vb + v);
      st4(b.x, vb + v, x0);
      st4(b.p, vb + v, make_float4(0.0f, 0.0f, 0.0f, x0.w));
      st4(b.v, vb + v, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
      b.drag[vb + v] = 0.0f;
    }
    if (tid < 12) b.tips[env * 12 + tid] = r.tips0[env * 12 + tid];
    if (tid < 12 && r.base0 && b.caps_base) b.caps_base[env * 12 + tid] = r.base0[env * 12 + tid];
    __syncthreads();
    if (b.bvh_lo && b.level_nodes && b.level_info) refit_env<kServeBlock>(b, e);
    return;
  }
  // step: move the four instrument points, then (optionally) reselect the grasp
  const float* a = r.action + 4 * (int64_t)i;
  if (tid < 12) {
    const int k = tid % 3;
    const float t0 = b.tips[env * 12 + tid];
    float t = t0 + a[k];
    if (r.lo) t = fmax_(t, r.lo[k]);
    if (r.hi) t = fmin_(t, r.hi[k]);
    b.tips[env * 12 + tid] = t;
    // capsule bases follow their tips (the instrument is translated rigidly)
    if (b.caps_base) b.caps_base[env * 12 + tid] += t - t0;
    if (tid >= 9) tip3[k] = t;
  }
  const float clamp = a[3];
  if (clamp < 0.0f) return;  // uniform per workgroup: no barrier is skipped by part of it
  __syncthreads();
  const f3 t3 = mk(tip3[0], tip3[1], tip3[2]);
  for (int v = tid; v < e.nv; v += kServeBlock) {
    float dr = 0.0f;
    if (clamp > 0.0f) dr = length(xyz(ld4(b.x, vb + v)) - t3) < radius ? 1.0f : 0.0f;
    b.drag[vb + v] = dr;
  }
}

// Block reduction in a fixed tree (deterministic for a given block size).
template <typename Op>
__device__ __forceinline__ float block_reduce(float v, float* sh, Op op) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = kServeBlock / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = op(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  const float out = sh[0];
  __syncthreads();
  return out;
}

__global__ __launch_bounds__(kServeBlock) void k_observe(ffsrl_batch b, const int32_t* envs, const float* x0,
                                                         float ground, float* out) {
  __shared__ float sh[kServeBlock];
  const int env = envs[blockIdx.x];
  const ffsrl_env_desc e = b.envs[env];
  const int64_t vb = e.vbase;
  const int tid = threadIdx.x;
  constexpr float kBig = 3.0e38f;
  float s[3] = {0, 0, 0}, mn[3] = {kBig, kBig, kBig}, mx[3] = {-kBig, -kBig, -kBig}, sv[3] = {0, 0, 0};
  float gs[3] = {0, 0, 0}, vmax = 0.0f, ng = 0.0f, nground = 0.0f, dmax = 0.0f, dsum = 0.0f;
  for (int v = tid; v < e.nv; v += kServeBlock) {
    const float4 x = ld4(b.x, vb + v), vel = ld4(b.v, vb + v), r0 = ld4(x0, vb + v);
    const float xs[3] = {x.x, x.y, x.z}, vs[3] = {vel.x, vel.y, vel.z};
    const bool grasped = b.drag[vb + v] != 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s[k] += xs[k];
      mn[k] = fmin_(mn[k], xs[k]);
      mx[k] = fmax_(mx[k], xs[k]);
      sv[k] += vs[k];
      if (grasped) gs[k] += xs[k];
    }
    vmax = fmax_(vmax, length(xyz(vel)));
    if (grasped) ng += 1.0f;
    if (x.y <= ground) nground += 1.0f;
    const float d = length(xyz(x) - xyz(r0));
    dmax = fmax_(dmax, d);
    dsum += d;
  }
  auto add = [](float a, float c) { return a + c; };
  auto fmin2 = [](float a, float c) { return fmin_(a, c); };
  auto fmax2 = [](float a, float c) { return fmax_(a, c); };
  float o[32];
  const float inv_n = e.nv > 0 ? 1.0f / (float)e.nv : 0.0f;
  for (int k = 0; k < 12; ++k) o[k] = b.tips[env * 12 + k];
  for (int k = 0; k < 3; ++k) {
    o[12 + k] = block_reduce(s[k], sh, add) * inv_n;
    o[15 + k] = block_reduce(mn[k], sh, fmin2);
    o[18 + k] = block_reduce(mx[k], sh, fmax2);
    o[21 + k] = block_reduce(sv[k], sh, add) * inv_n;
  }
  o[24] = block_reduce(vmax, sh, fmax2);
  o[25] = block_reduce(ng, sh, add);
  for (int k = 0; k < 3; ++k) {
    const float g = block_reduce(gs[k], sh, add);
    o[26 + k] = o[25] > 0.0f ? g / o[25] : 0.0f;
  }
  o[29] = block_reduce(nground, sh, add);
  o[30] = block_reduce(dmax, sh, fmax2);
  o[31] = block_reduce(dsum, sh, add) * inv_n;
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < 32; ++k) out[(int64_t)env * 32 + k] = o[k];
  }
}

}  // namespace ffsrl

using namespace ffsrl;

extern "C" {

int ffsrl_apply_requests(const ffsrl_batch* b, const ffsrl_requests* r, float radius, void* stream) {
  if (!b || !r || !b->envs || !b->x || !b->p || !b->v || !b->drag || !b->tips)
    return fail(FFSRL_E_ARG, "apply_requests: NULL batch or state");
  if (r->n < 0 || r->n > b->n_env) return fail(FFSRL_E_ARG, "apply_requests: %d requests for %d envs", r->n, b->n_env);
  if (r->n == 0) return 0;
  if (!r->envs || !r->kind || !r->action || !r->x0 || !r->tips0)
    return fail(FFSRL_E_ARG, "apply_requests: NULL request array");
  hipLaunchKernelGGL(k_apply_requests, dim3(r->n), dim3(kServeBlock), 0, as_stream(stream), *b, *r, radius);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_observe_envs(const ffsrl_batch* b, const int32_t* envs, int32_t n, const float* x0, float ground, float* out,
                       void* stream) {
  if (!b || !b->envs || !b->x || !b->v || !b->drag || !b->tips) return fail(FFSRL_E_ARG, "observe: NULL batch");
  if (n < 0 || n > b->n_env) return fail(FFSRL_E_ARG, "observe: %d envs listed, batch has %d", n, b->n_env);
  if (n == 0) return 0;
  if (!envs || !x0 || !out) return fail(FFSRL_E_ARG, "observe: NULL argument");
  hipLaunchKernelGGL(k_observe, dim3(n), dim3(kServeBlock), 0, as_stream(stream), *b, envs, x0, ground, out);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
