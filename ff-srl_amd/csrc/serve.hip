// Request kernels of the serving loop: apply a tick's reset / step requests to
// the listed environments, and reduce each listed environment to its
// observation scalars.  Both run one 1024-thread workgroup per listed
// environment, so a tick of mixed requests is two launches (plus the frame
// and the render).
//
// Reset follows resetFixedModel (modelDO.py:1589-1622): positions restored,
// velocities and grasp weights zeroed, instrument points restored; the env's
// BVH is refit in the same workgroup so the reset observation renders the
// restored mesh.  A step moves the instrument points by the request's delta,
// less the components that would take the drag point (point 3) out of the
// workspace box, as transfromCartesianActionsKernelInWorkspace drops them
// (modelDO.py:97-154); the whole instrument moves by the same delta.  When
// the request opens or closes the clamp,
// reselects the grasped vertices like updateClampsDragKernel
// (modelDO.py:247-279): drag = 1 for vertices within `radius` of the drag
// point tip[env*4+3] when closed, 0 when open.
#include "common.hpp"
#include "refit.hpp"

namespace ffsrl {

constexpr int kServeBlock = 1024;
constexpr int kServeWaves = kServeBlock / 64;

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
    if (tid == 0 && b.status) b.status[env] = 0;  // a new episode
    __syncthreads();
    if (b.bvh_lo && b.level_nodes && b.level_info) refit_env<kServeBlock>(b, e);
    return;
  }
  // step: move the four instrument points, then (optionally) reselect the grasp
  const float* a = r.action + 4 * (int64_t)i;
  __shared__ float delta[3];
  if (tid < 3) {
    // a component that would take the drag point out of the workspace is dropped
    const float d = a[tid], moved = b.tips[env * 12 + 9 + tid] + d;
    delta[tid] = ((r.lo && moved < r.lo[tid]) || (r.hi && moved > r.hi[tid])) ? 0.0f : d;
  }
  __syncthreads();
  if (tid < 12) {
    const int k = tid % 3;
    const float t0 = b.tips[env * 12 + tid];
    const float t = t0 + delta[k];
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

// The per-vertex quantities an observation reduces, in three groups:
//   sums    0-2 x, 3-5 v, 6-8 grasped x, 9 grasped count, 10 on-ground
//           count, 11 displacement from the initial position;
//   minima  12-14 x;
//   maxima  15-17 x, 18 |v|, 19 displacement.
constexpr int kObsSums = 12, kObsMins = 3, kObsQ = 20;

__device__ __forceinline__ float obs_op(int k, float a, float c) {
  return k < kObsSums ? a + c : k < kObsSums + kObsMins ? fmin_(a, c) : fmax_(a, c);
}

// One vertex into the thread's partial quantities.
__device__ __forceinline__ void observe_vertex(float q[kObsQ], float4 x, float4 vel, float4 r0, bool grasped,
                                               float ground) {
  const float xs[3] = {x.x, x.y, x.z}, vs[3] = {vel.x, vel.y, vel.z};
  const float d = length(xyz(x) - xyz(r0));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    q[k] += xs[k];
    q[3 + k] += vs[k];
    if (grasped) q[6 + k] += xs[k];
    q[12 + k] = fmin_(q[12 + k], xs[k]);
    q[15 + k] = fmax_(q[15 + k], xs[k]);
  }
  if (grasped) q[9] += 1.0f;
  if (x.y <= ground) q[10] += 1.0f;
  q[11] += d;
  q[18] = fmax_(q[18], length(xyz(vel)));
  q[19] = fmax_(q[19], d);
}

// Each thread takes the vertices v = tid + k * kServeBlock, four loads in
// flight per step; every quantity is then reduced across the wavefront with
// xor shuffles and across the wavefronts through LDS, in a fixed order, so an
// env's observation does not depend on what else the launch holds.
__global__ __launch_bounds__(kServeBlock) void k_observe(ffsrl_batch b, const int32_t* envs, const float* x0,
                                                         float ground, float* out) {
  __shared__ float sh[kObsQ][kServeWaves];
  const int env = envs[blockIdx.x];
  const ffsrl_env_desc e = b.envs[env];
  const int64_t vb = e.vbase;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr float kBig = 3.0e38f;
  float q[kObsQ];
#pragma unroll
  for (int k = 0; k < kObsQ; ++k) q[k] = k < kObsSums ? 0.0f : k < kObsSums + kObsMins ? kBig : -kBig;
  q[18] = 0.0f;  // speeds and displacements are >= 0
  q[19] = 0.0f;
  constexpr int U = 4;
  for (int v0 = tid; v0 < e.nv; v0 += U * kServeBlock) {
    float4 x[U], vel[U], r0[U];
    float dr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int v = v0 + u * kServeBlock;
      if (v < e.nv) {
        x[u] = ld4(b.x, vb + v);
        vel[u] = ld4(b.v, vb + v);
        r0[u] = ld4(x0, vb + v);
        dr[u] = b.drag[vb + v];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (v0 + u * kServeBlock < e.nv) observe_vertex(q, x[u], vel[u], r0[u], dr[u] != 0.0f, ground);
  }
#pragma unroll
  for (int k = 0; k < kObsQ; ++k) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) q[k] = obs_op(k, q[k], __shfl_xor(q[k], m));
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kObsQ; ++k) sh[k][wave] = q[k];
  }
  __syncthreads();
  if (tid == 0) {
    float r[kObsQ];
#pragma unroll
    for (int k = 0; k < kObsQ; ++k) r[k] = sh[k][0];
    for (int w = 1; w < kServeWaves; ++w) {
#pragma unroll
      for (int k = 0; k < kObsQ; ++k) r[k] = obs_op(k, r[k], sh[k][w]);
    }
    const float inv_n = e.nv > 0 ? 1.0f / (float)e.nv : 0.0f;
    float* o = out + (int64_t)env * 32;
    for (int k = 0; k < 12; ++k) o[k] = b.tips[env * 12 + k];  // the four instrument points
    for (int k = 0; k < 3; ++k) {
      o[12 + k] = r[k] * inv_n;                           // centroid
      o[15 + k] = r[12 + k];                              // bounding box min
      o[18 + k] = r[15 + k];                              // bounding box max
      o[21 + k] = r[3 + k] * inv_n;                       // mean velocity
      o[26 + k] = r[9] > 0.0f ? r[6 + k] / r[9] : 0.0f;  // centroid of the grasped vertices
    }
    o[24] = r[18];          // max speed
    o[25] = r[9];           // grasped vertices
    o[29] = r[10];          // vertices on the ground
    o[30] = r[19];          // max displacement from the initial positions
    o[31] = r[11] * inv_n;  // mean displacement
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
