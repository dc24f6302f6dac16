// Instrument kinematics: the seven transforms of each env's laparoscope, the
// jaw state machine, the remote-centre-of-motion Cartesian actions and the
// capsule points derived from them (modelDO.py:55-289, 1331-1404, 1570-1587).
//
// One thread per env: an env's transforms depend on each other in a fixed
// order, and the work is a few hundred flops per env.  Transforms are kept
// in the reference's storage (transposed, row-vector convention); the
// arithmetic below works on column-vector matrices m[row][col] and writes
// them back transposed.  Every matrix product sums k = 0..3 left to right,
// as a serial loop does.

#include "common.hpp"

namespace ffsrl {
namespace {

constexpr float kPi = 3.14159265358979323846f;  // wp.constant(math.pi) as float
constexpr int kRod = 0, kRcm = 3, kLeftPivot = 4, kRightPivot = 5;
constexpr float kJawDrag = 3.0f, kJawMax = 85.0f, kJawMin = 40.0f;  // modelDO.py:20-22 (degrees)

struct M4 {
  float m[4][4];
};

__device__ inline M4 load_xf(const float* xf, int id) {  // stored transposed
  M4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) r.m[i][j] = xf[id * 16 + j * 4 + i];
  return r;
}

__device__ inline void store_xf(float* xf, int id, const M4& a) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) xf[id * 16 + j * 4 + i] = a.m[i][j];
}

__device__ inline M4 transpose(const M4& a) {
  M4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) r.m[i][j] = a.m[j][i];
  return r;
}

__device__ inline M4 mul(const M4& a, const M4& b) {
  M4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) s = s + a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}

// Inverse by cofactors over 2x2 sub-determinants of the first two and the
// last two rows; all-zero when the determinant is 0.
__device__ inline M4 inverse(const M4& a) {
  const float(*m)[4] = a.m;
  const float s0 = m[0][0] * m[1][1] - m[1][0] * m[0][1];
  const float s1 = m[0][0] * m[1][2] - m[1][0] * m[0][2];
  const float s2 = m[0][0] * m[1][3] - m[1][0] * m[0][3];
  const float s3 = m[0][1] * m[1][2] - m[1][1] * m[0][2];
  const float s4 = m[0][1] * m[1][3] - m[1][1] * m[0][3];
  const float s5 = m[0][2] * m[1][3] - m[1][2] * m[0][3];
  const float c5 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
  const float c4 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
  const float c3 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
  const float c2 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
  const float c1 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
  const float c0 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
  const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  M4 r;
  if (det == 0.0f) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) r.m[i][j] = 0.0f;
    return r;
  }
  const float id = 1.0f / det;
  r.m[0][0] = (m[1][1] * c5 - m[1][2] * c4 + m[1][3] * c3) * id;
  r.m[0][1] = (-m[0][1] * c5 + m[0][2] * c4 - m[0][3] * c3) * id;
  r.m[0][2] = (m[3][1] * s5 - m[3][2] * s4 + m[3][3] * s3) * id;
  r.m[0][3] = (-m[2][1] * s5 + m[2][2] * s4 - m[2][3] * s3) * id;
  r.m[1][0] = (-m[1][0] * c5 + m[1][2] * c2 - m[1][3] * c1) * id;
  r.m[1][1] = (m[0][0] * c5 - m[0][2] * c2 + m[0][3] * c1) * id;
  r.m[1][2] = (-m[3][0] * s5 + m[3][2] * s2 - m[3][3] * s1) * id;
  r.m[1][3] = (m[2][0] * s5 - m[2][2] * s2 + m[2][3] * s1) * id;
  r.m[2][0] = (m[1][0] * c4 - m[1][1] * c2 + m[1][3] * c0) * id;
  r.m[2][1] = (-m[0][0] * c4 + m[0][1] * c2 - m[0][3] * c0) * id;
  r.m[2][2] = (m[3][0] * s4 - m[3][1] * s2 + m[3][3] * s0) * id;
  r.m[2][3] = (-m[2][0] * s4 + m[2][1] * s2 - m[2][3] * s0) * id;
  r.m[3][0] = (-m[1][0] * c3 + m[1][1] * c1 - m[1][2] * c0) * id;
  r.m[3][1] = (m[0][0] * c3 - m[0][1] * c1 + m[0][2] * c0) * id;
  r.m[3][2] = (-m[3][0] * s3 + m[3][1] * s1 - m[3][2] * s0) * id;
  r.m[3][3] = (m[2][0] * s3 - m[2][1] * s1 + m[2][2] * s0) * id;
  return r;
}

// (m * (p, 1)).xyz, each row summed left to right
__device__ inline f3 transform_point(const M4& a, f3 p) {
  float o[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = ((a.m[i][0] * p.x + a.m[i][1] * p.y) + a.m[i][2] * p.z) + a.m[i][3];
  return mk(o[0], o[1], o[2]);
}

struct Quat {
  float x, y, z, w;
};

__device__ inline Quat quat_axis_angle(f3 axis, float angle) {
  const float h = angle * 0.5f;
  const float s = sinf(h);
  return Quat{axis.x * s, axis.y * s, axis.z * s, cosf(h)};
}

// q x q^-1 = x (2w^2 - 1) + 2 w (v x x) + 2 (v.x) v
__device__ inline f3 quat_rotate(Quat q, f3 x) {
  const f3 v = mk(q.x, q.y, q.z);
  const float c = 2.0f * q.w * q.w - 1.0f;
  const float d = 2.0f * dot(v, x);
  return (x * c + cross(v, x) * (2.0f * q.w)) + v * d;
}

// rotation q (columns: the rotated axes) then translation t, unit scale
__device__ inline M4 rigid(f3 t, Quat q) {
  const f3 c0 = quat_rotate(q, mk(1.0f, 0.0f, 0.0f));
  const f3 c1 = quat_rotate(q, mk(0.0f, 1.0f, 0.0f));
  const f3 c2 = quat_rotate(q, mk(0.0f, 0.0f, 1.0f));
  return M4{{{c0.x, c1.x, c2.x, t.x}, {c0.y, c1.y, c2.y, t.y}, {c0.z, c1.z, c2.z, t.z}, {0.0f, 0.0f, 0.0f, 1.0f}}};
}

// getTransformationMatrix (modelDO.py:396-416): T * (Rx * (Ry * Rz))
__device__ inline M4 transformation(float sx, float sy, float sz, float ax, float ay, float az) {
  const f3 zero = mk(0.0f, 0.0f, 0.0f);
  const M4 shift = rigid(mk(sx, sy, sz), quat_axis_angle(zero, 0.0f));
  const M4 rx = rigid(zero, quat_axis_angle(mk(1.0f, 0.0f, 0.0f), ax));
  const M4 ry = rigid(zero, quat_axis_angle(mk(0.0f, 1.0f, 0.0f), ay));
  const M4 rz = rigid(zero, quat_axis_angle(mk(0.0f, 0.0f, 1.0f), az));
  return mul(shift, mul(rx, mul(ry, rz)));
}

// applyTransformationToChild (modelDO.py:386-394): C' = P * (M * (P^-1 * C))
__device__ inline void carry_child(float* xf, const M4& mat, int parent, int child) {
  const M4 P = load_xf(xf, parent);
  const M4 Pi = inverse(P);
  const M4 C = load_xf(xf, child);
  store_xf(xf, child, mul(P, mul(mat, mul(Pi, C))));
}

// applyTransformationToChildren (modelDO.py:364-384) for env-local kind k
__device__ inline void carry_children(float* xf, const M4& mat, int base, int kind) {
  if (kind == kLeftPivot) carry_child(xf, mat, base + kind, base + 1);
  if (kind == kRightPivot) carry_child(xf, mat, base + kind, base + 2);
  if (kind == kRcm)
    for (int i = 0; i < 7; ++i)
      if (i != kind) carry_child(xf, mat, base + kind, base + i);
  if (kind == kRod)
    for (int i = 1; i < 7; ++i)
      if (i != kRcm) carry_child(xf, mat, base + kind, base + i);
}

// transformFunc (modelDO.py:291-299): children first (old parent), then
// the parent becomes X * M
__device__ inline void transform_node(float* xf, int base, int kind, const M4& X, const M4& mat) {
  carry_children(xf, mat, base, kind);
  store_xf(xf, base + kind, mul(X, mat));
}

// rotateClampsFunc (modelDO.py:301-362)
__device__ inline void rotate_jaws(float* xf, int32_t* dch, int32_t* dchu, float action, int env) {
  const int base = env * 7;
  float diff = action * 180.0f / kPi;
  if (diff == 0.0f) return;
  const M4 rodInv = inverse(load_xf(xf, base + kRod));
  const M4 left = load_xf(xf, base + kLeftPivot);
  const M4 right = load_xf(xf, base + kRightPivot);
  const M4 rightLocal = mul(rodInv, right);
  const float angle = asinf(rightLocal.m[2][1]) * 180.0f / kPi;
  bool apply = false;
  const int drag = env * 3;
  if (dch[drag] == 1 && dchu[drag] == 1) dchu[drag] = 0;
  if (action > 0.0f) {
    if (angle < kJawMax) {
      if (angle + diff > kJawMax) diff = kJawMax - angle;
      apply = true;
    }
    if (angle > kJawDrag + kJawMin && dch[drag] == 1) {
      dch[drag] = 0;
      dchu[drag] = 1;
    }
  } else {
    if (angle > kJawMin) {
      if (angle + diff < kJawMin) diff = kJawMin - angle;
      apply = true;
    }
    if (angle < kJawDrag + kJawMin && dch[drag] == 0) {
      dch[drag] = 1;
      dchu[drag] = 1;
    }
  }
  if (!apply) return;
  const float rad = diff * kPi / 180.0f;
  const M4 rot = transformation(0.0f, 0.0f, 0.0f, rad, 0.0f, 0.0f);
  const M4 rotInv = transformation(0.0f, 0.0f, 0.0f, -rad, 0.0f, 0.0f);
  // the pivot is updated before its jaw here (the parent of the carry is the new one)
  store_xf(xf, base + kLeftPivot, mul(left, rotInv));
  carry_children(xf, rotInv, base, kLeftPivot);
  store_xf(xf, base + kRightPivot, mul(right, rot));
  carry_children(xf, rot, base, kRightPivot);
}

// updateLaparoscopeKernel (modelDO.py:218-245)
__device__ inline void update_points(const ffsrl_batch& b, const ffsrl_instrument& ins, int env) {
  for (int k = 0; k < 4; ++k) {
    M4 X;
    f3 lo, hi;
    if (k < 3) {
      X = load_xf(ins.xform, env * 7 + k);
      const float h = ins.height[env * 4 + k];
      lo = mk(0.0f, 0.0f, 0.0f - h / 1.85f);
      hi = mk(0.0f, 0.0f, 0.0f + h / 1.85f);
    } else {
      X = load_xf(ins.xform, env * 7);
      const float rod = ins.height[env * 4], jaw = ins.height[env * 4 + 2];
      lo = mk(0.0f, 0.0f, 0.0f - rod * 0.5f - jaw * 1.0f);
      hi = mk(0.0f, 0.0f, 0.0f + rod * 0.5f + jaw * 1.0f);
    }
    const f3 pb = transform_point(X, lo), pt = transform_point(X, hi);
    const int c = env * 4 + k;
    if (b.caps_base) {
      b.caps_base[3 * c] = pb.x;
      b.caps_base[3 * c + 1] = pb.y;
      b.caps_base[3 * c + 2] = pb.z;
    }
    b.tips[3 * c] = pt.x;
    b.tips[3 * c + 1] = pt.y;
    b.tips[3 * c + 2] = pt.z;
  }
}

__global__ void k_instrument_actions(ffsrl_batch b, ffsrl_instrument ins, const float* actions) {
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= ins.n_env) return;
  const float* a = actions + 5 * env;
  const int base = env * 7;
  float* xf = ins.xform;
  // applyActionsKernel (modelDO.py:175-199): its rod, rotation-centre and jaw
  // threads of this env, in thread order
  transform_node(xf, base, kRod, load_xf(xf, base + kRod), transformation(0.0f, 0.0f, a[0], 0.0f, 0.0f, a[1]));
  transform_node(xf, base, kRcm, load_xf(xf, base + kRcm), transformation(0.0f, 0.0f, 0.0f, a[2], a[3], 0.0f));
  rotate_jaws(xf, ins.drag_cut_heat, ins.drag_cut_heat_update, a[4], env);
  update_points(b, ins, env);
}

// transfromCartesianActionsKernel[InWorkspace] (modelDO.py:55-154) for a
// nonzero action
__device__ inline void cartesian_env(const ffsrl_batch& b, const ffsrl_instrument& ins, const float* a,
                                     const float* lo, const float* hi, int env) {
  const int base = env * 7;
  float* xf = ins.xform;
  const int d = (env * 4 + 3) * 3;
  const f3 cur = mk(b.tips[d], b.tips[d + 1], b.tips[d + 2]);
  float av[3] = {a[0], a[1], a[2]};
  if (lo && hi) {
    int viol = 0;
    const float nw[3] = {cur.x + av[0], cur.y + av[1], cur.z + av[2]};
    for (int c = 0; c < 3; ++c)
      if (nw[c] < lo[c] || nw[c] > hi[c]) {
        av[c] = 0.0f;
        viol = 1;
      }
    if (ins.violation) ins.violation[env] = viol;
  }
  const M4 rodX = load_xf(xf, base + kRod);
  const M4 rcS = transpose(load_xf(xf, base + kRcm));  // the stored (transposed) matrix
  const f3 rel = transform_point(transpose(inverse(rcS)), cur);
  const f3 delta = transform_point(rcS, mk(av[0], av[1], av[2]));  // R^T applied to the action
  const f3 nrel = rel + delta;
  const float r0 = length(rel), r1 = length(nrel);
  const f3 vc = cross(rel, nrel);
  const float sin_t = length(vc) / (length(rel) * length(nrel));
  const f3 axis = normalize(vc);
  const float angle = asinf(sin_t);
  const Quat q = quat_axis_angle(axis, angle);
  transform_node(xf, base, kRod, rodX, transformation(0.0f, 0.0f, r1 - r0, 0.0f, 0.0f, 0.0f));
  transform_node(xf, base, kRcm, transpose(rcS), rigid(mk(0.0f, 0.0f, 0.0f), q));
}

// the action kernel (a zero action leaves the env, flags included), then
// updateLaparoscopeKernel for every env
__global__ void k_instrument_cartesian(ffsrl_batch b, ffsrl_instrument ins, const float* actions, const float* lo,
                                       const float* hi) {
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= ins.n_env) return;
  const float* a = actions + 3 * env;
  if (!(a[0] == 0.0f && a[1] == 0.0f && a[2] == 0.0f)) cartesian_env(b, ins, a, lo, hi, env);
  update_points(b, ins, env);
}

// setRotationCenterTransformKernel (modelDO.py:156-173)
__global__ void k_instrument_set_rcm(ffsrl_batch b, ffsrl_instrument ins, const float* t, const float* r) {
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= ins.n_env) return;
  const int base = env * 7;
  const M4 X = load_xf(ins.xform, base + kRcm);
  const M4 target = transformation(t[3 * env], t[3 * env + 1], t[3 * env + 2], r[3 * env], r[3 * env + 1],
                                   r[3 * env + 2]);
  transform_node(ins.xform, base, kRcm, X, mul(inverse(X), target));
  update_points(b, ins, env);
}

__global__ void k_instrument_update(ffsrl_batch b, ffsrl_instrument ins) {
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= ins.n_env) return;
  update_points(b, ins, env);
}

// updateClampsDragKernel (modelDO.py:247-279) over all envs in one launch:
// block (x, env) covers vertices x*256.. of env
__global__ void k_instrument_grasp(ffsrl_batch b, float radius, const int32_t* state, const int32_t* update) {
  const int env = blockIdx.y;
  const ffsrl_env_desc e = b.envs[env];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= e.nv || update[env * 3] == 0) return;
  const int64_t g = (int64_t)e.vbase + i;
  float dr = 0.0f;
  if (state[env * 3] == 1) {
    const float* t = b.tips + (env * 4 + 3) * 3;
    dr = length(xyz(ld4(b.p, g)) - mk(t[0], t[1], t[2])) < radius ? 1.0f : 0.0f;
  }
  b.drag[g] = dr;
}

int check(const ffsrl_batch* b, const ffsrl_instrument* ins) {
  if (!b || !ins) return fail(FFSRL_E_ARG, "instrument: NULL batch or instrument");
  if (!ins->xform || !ins->height || !b->tips) return fail(FFSRL_E_ARG, "instrument: NULL transforms, heights or tips");
  if (ins->n_env != b->n_env) return fail(FFSRL_E_ARG, "instrument has %d envs, batch %d", ins->n_env, b->n_env);
  return 0;
}

dim3 env_grid(int n) { return dim3((n + 63) / 64); }

}  // namespace
}  // namespace ffsrl

using namespace ffsrl;

extern "C" {

int ffsrl_instrument_actions(const ffsrl_batch* b, const ffsrl_instrument* ins, const float* actions,
                             float lookup_radius, void* stream) {
  int rc = check(b, ins);
  if (rc) return rc;
  if (!actions || !ins->drag_cut_heat || !ins->drag_cut_heat_update)
    return fail(FFSRL_E_ARG, "instrument actions: NULL actions or clamp state");
  if (lookup_radius >= 0.0f && (!b->envs || !b->p || !b->drag))
    return fail(FFSRL_E_ARG, "instrument actions: grasp selection needs envs, positions and drag");
  if (lookup_radius >= 0.0f && b->n_env > 65535)
    return fail(FFSRL_E_ARG, "instrument actions: grasp selection launches one grid row per env (<= 65535)");
  if (ins->n_env == 0) return 0;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_instrument_actions, env_grid(ins->n_env), dim3(64), 0, st, *b, *ins, actions);
  if (lookup_radius >= 0.0f && b->max_nv > 0)
    hipLaunchKernelGGL(k_instrument_grasp, dim3((b->max_nv + 255) / 256, b->n_env), dim3(256), 0, st, *b,
                       lookup_radius, ins->drag_cut_heat, ins->drag_cut_heat_update);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_instrument_cartesian(const ffsrl_batch* b, const ffsrl_instrument* ins, const float* actions,
                               const float* lo, const float* hi, void* stream) {
  int rc = check(b, ins);
  if (rc) return rc;
  if (!actions) return fail(FFSRL_E_ARG, "instrument cartesian: NULL actions");
  if ((lo == nullptr) != (hi == nullptr)) return fail(FFSRL_E_ARG, "instrument cartesian: give both lo and hi");
  if (ins->n_env == 0) return 0;
  hipLaunchKernelGGL(k_instrument_cartesian, env_grid(ins->n_env), dim3(64), 0, as_stream(stream), *b, *ins,
                     actions, lo, hi);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_instrument_set_rcm(const ffsrl_batch* b, const ffsrl_instrument* ins, const float* translation,
                             const float* rotation, void* stream) {
  int rc = check(b, ins);
  if (rc) return rc;
  if (!translation || !rotation) return fail(FFSRL_E_ARG, "instrument rotation centre: NULL translation or rotation");
  if (ins->n_env == 0) return 0;
  hipLaunchKernelGGL(k_instrument_set_rcm, env_grid(ins->n_env), dim3(64), 0, as_stream(stream), *b, *ins,
                     translation, rotation);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_instrument_update(const ffsrl_batch* b, const ffsrl_instrument* ins, void* stream) {
  int rc = check(b, ins);
  if (rc) return rc;
  if (ins->n_env == 0) return 0;
  hipLaunchKernelGGL(k_instrument_update, env_grid(ins->n_env), dim3(64), 0, as_stream(stream), *b, *ins);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
