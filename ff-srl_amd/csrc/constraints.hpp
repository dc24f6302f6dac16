// Per-constraint projection math, shared by the batched (LDS) and plain
// (global-atomic) kernels.  Each function follows the reference kernel's
// expression; positions arrive as float4 (x, y, z, inverse mass).
//
// Arithmetic policy Ar<FAST>:
//   FAST = false: IEEE division and square root, the reference's operation
//                 order (Neo-Hookean: up to the order of additions; the
//                 plain form);
//   FAST = true:  hardware reciprocal / square root / reciprocal square root
//                 (<= 1 ulp), multiplications by reciprocals, squared lengths
//                 taken as dot products (the batched frame).
// Both are within the positional tolerance of the reference; the choice only
// moves the last bits.  This file is compiled with multiply-add contraction
// (no bit-exact result depends on it: the refit is min/max only).
//
// `scale` (a power of two, 1 in the plain form) multiplies the returned
// deltas: the batched frames pass 2^-dp_exp and get their fixed-point
// quanta directly.  It is applied to the constraint's scalar multiplier
// (dlambda); scaling by a power of two commutes with rounding, so the
// deltas are exactly scale times the unscaled ones (lambda itself stays
// unscaled).
#pragma once

#include "common.hpp"

namespace ffsrl {

template <bool FAST>
struct Ar;

template <>
struct Ar<false> {
  static __device__ __forceinline__ float div(float a, float b) { return a / b; }
  static __device__ __forceinline__ float sqrt(float x) { return sqrtf(x); }
  static __device__ __forceinline__ f3 vdiv(f3 a, float s) { return a / s; }
  static __device__ __forceinline__ float sq_len(f3 a) {  // length(a) * length(a)
    float l = length(a);
    return l * l;
  }
};

template <>
struct Ar<true> {
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float div(float a, float b) { return a * rcp(b); }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ f3 vdiv(f3 a, float s) { return a * rcp(s); }
  static __device__ __forceinline__ float sq_len(f3 a) { return dot(a, a); }
};

// distanceConstraints, PBD form (integrator.py:65-109), all edges active.
// On success returns 1 and e = edgeDP; the caller applies dP[a] -= e*wa and
// dP[b] += e*wb and counts both ends.
template <bool FAST>
__device__ __forceinline__ int distance_pbd(float4 A, float4 B, float rest, float kS, f3* e, float scale = 1.0f) {
  using M = Ar<FAST>;
  if (rest < kFeps) return 0;
  f3 dir = xyz(B) - xyz(A);
  float len = M::sqrt(dot(dir, dir));
  if (len < kFeps) return 0;
  float w = A.w + B.w;
  if (w < kFeps) return 0;
  float C = len - rest;
  if (FAST) {
    *e = dir * ((M::div(-C * kS, w) * M::div(1.0f, len)) * scale);
  } else {
    *e = ((dir / len) * (-C)) * kS / w;
  }
  return 1;
}

// distanceConstraintsXPBD (integrator.py:191-240); alpha = compliance/dt^2.
// *lam is read and updated in place.
template <bool FAST>
__device__ __forceinline__ int distance_xpbd(float4 A, float4 B, float rest, float alpha, float* lam, f3* e,
                                             float scale = 1.0f) {
  using M = Ar<FAST>;
  if (rest < kFeps) return 0;
  f3 dir = xyz(B) - xyz(A);
  float len = M::sqrt(dot(dir, dir));
  if (len < kFeps) return 0;
  float w = A.w + B.w;
  if (w < kFeps) return 0;
  float C = len - rest;
  float dl = M::div(-(C + alpha * (*lam)), w + alpha);
  *lam = *lam + dl;
  *e = M::vdiv(dir, len) * (dl * scale);  // normalize(dir) * dlambda
  return 1;
}

struct VolTerms {
  f3 q[4];
  float six_vol, w;
};

// q vectors (q1 = p2 x p3, q2 = p3 x p1, q3 = p1 x p2, q0 = -q1-q2-q3), the
// signed 6V and w = sum m_i |q_i|^2 of the volume kernels (integrator.py:474-490).
__device__ __forceinline__ VolTerms vol_terms(const float4 P[4]) {
  VolTerms t;
  f3 a = xyz(P[0]);
  f3 p1 = xyz(P[1]) - a, p2 = xyz(P[2]) - a, p3 = xyz(P[3]) - a;
  t.q[2] = cross(p3, p1);
  t.q[1] = cross(p2, p3);
  t.q[3] = cross(p1, p2);
  t.q[0] = ((-t.q[1]) - t.q[2]) - t.q[3];
  t.six_vol = dot(cross(p1, p2), p3);
  t.w = P[0].w * dot(t.q[0], t.q[0]) + P[1].w * dot(t.q[1], t.q[1]) + P[2].w * dot(t.q[2], t.q[2]) +
        P[3].w * dot(t.q[3], t.q[3]);
  return t;
}

// volumeConstraints, PBD form (integrator.py:444-504): |V|, C = V0 - |V|.
template <bool FAST>
__device__ __forceinline__ int volume_pbd(const float4 P[4], float rest, float kS, f3 d[4], float scale = 1.0f) {
  VolTerms t = vol_terms(P);
  float volume = FAST ? fabsf(t.six_vol) * (1.0f / 6.0f) : fabsf(t.six_vol) / 6.0f;
  if (fabsf(t.w) < kFeps) return 0;
  float C = rest - volume;
  const float dl = Ar<FAST>::div(-kS * C, t.w) * scale;
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = (t.q[i] * dl) * P[i].w;
  return 1;
}

// volumeConstraintsXPBD (integrator.py:507-576): signed V, C = V - V0.
template <bool FAST>
__device__ __forceinline__ int volume_xpbd(const float4 P[4], float rest, float alpha, float* lam, f3 d[4],
                                           float scale = 1.0f) {
  VolTerms t = vol_terms(P);
  float volume = FAST ? t.six_vol * (1.0f / 6.0f) : t.six_vol / 6.0f;
  if (fabsf(t.w) < kFeps) return 0;
  float C = volume - rest;
  float dl = Ar<FAST>::div(-(C + alpha * (*lam)), t.w + alpha);
  *lam = *lam + dl;
  const float ds = dl * scale;
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = (t.q[i] * ds) * P[i].w;
  return 1;
}

// Constants of neoHookeanConstraints that do not depend on the tet.
struct NHConst {
  float gamma;     // invD / invV = (1/devC) / (1/volC)
  float alpha_h;   // volC / dT / dT  (divided by the rest volume per tet)
  float alpha_d;   // devC / dT / dT
};

__host__ __device__ inline NHConst nh_const(float volC, float devC, float dT) {
  NHConst c;
  float invV = 1.0f / volC, invD = 1.0f / devC;
  c.gamma = invD / invV;
  c.alpha_h = volC / dT / dT;
  c.alpha_d = devC / dT / dT;
  return c;
}

// M = row-major 3x3; F = Ds * M with Ds = columns (p1, p2, p3).
// neoHookeanConstraints (integrator.py:578-698): hydrostatic det F - 1 - gamma
// with gradient rows of transpose(mat33(F1xF2, F2xF0, F0xF1) * M^T), then
// deviatoric |F|_F with gradient rows of transpose(F * M^T / |F|); both
// deltas from the same positions; lambda is not carried.  The gradient of
// each vertex is formed when it is used (weights, then the summed delta)
// instead of as two 3x3 matrices, which keeps the batched frame within its
// 128 registers; the arithmetic is the reference's up to the order of the
// additions.
template <bool FAST>
__device__ __forceinline__ int neohookean(const float4 P[4], float rest, const float M[9], const NHConst& k,
                                          f3 d[4], float scale = 1.0f) {
  using A = Ar<FAST>;
  const f3 a = xyz(P[0]);
  const f3 p1 = xyz(P[1]) - a, p2 = xyz(P[2]) - a, p3 = xyz(P[3]) - a;
  const float m[4] = {P[0].w, P[1].w, P[2].w, P[3].w};
  // rows of F = Ds * M (Ds columns p1, p2, p3)
  f3 F[3];
  F[0] = mk(p1.x * M[0] + p2.x * M[3] + p3.x * M[6], p1.x * M[1] + p2.x * M[4] + p3.x * M[7],
            p1.x * M[2] + p2.x * M[5] + p3.x * M[8]);
  F[1] = mk(p1.y * M[0] + p2.y * M[3] + p3.y * M[6], p1.y * M[1] + p2.y * M[4] + p3.y * M[7],
            p1.y * M[2] + p2.y * M[5] + p3.y * M[8]);
  F[2] = mk(p1.z * M[0] + p2.z * M[3] + p3.z * M[6], p1.z * M[1] + p2.z * M[4] + p3.z * M[7],
            p1.z * M[2] + p2.z * M[5] + p3.z * M[8]);
  const f3 c1 = cross(F[1], F[2]), c2 = cross(F[2], F[0]), c3 = cross(F[0], F[1]);
  // gradient column of vertex j+1: hydrostatic g_j = c1 M[3j] + c2 M[3j+1] + c3 M[3j+2],
  // deviatoric h_j = (F0 . M_j, F1 . M_j, F2 . M_j) / |F|; vertex 0 takes minus their sum
  auto gcol = [&](int j) { return (c1 * M[3 * j] + c2 * M[3 * j + 1]) + c3 * M[3 * j + 2]; };
  auto hcol = [&](int j) {
    const f3 r = mk(M[3 * j], M[3 * j + 1], M[3 * j + 2]);
    return mk(dot(F[0], r), dot(F[1], r), dot(F[2], r));
  };
  f3 s = mk(0.0f, 0.0f, 0.0f);
  float w = 0.0f;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const f3 g = gcol(j);
    w = w + A::sq_len(g) * m[j + 1];
    s = s + g;
  }
  w = w + A::sq_len(s) * m[0];
  if (w < kFeps) return 0;
  const float C_h = dot(F[0], c1) - 1.0f - k.gamma;
  const float inv_rest = FAST ? __builtin_amdgcn_rcpf(rest) : 0.0f;
  const float dl_h = A::div(-C_h, w + (FAST ? k.alpha_h * inv_rest : k.alpha_h / rest)) * scale;
  float rsi, C_d;
  if (FAST) {
    const float r2 = (dot(F[0], F[0]) + dot(F[1], F[1])) + dot(F[2], F[2]);
    rsi = __builtin_amdgcn_rsqf(r2);
    C_d = r2 * rsi;
  } else {
    const float l0 = length(F[0]), l1 = length(F[1]), l2 = length(F[2]);
    C_d = sqrtf(l0 * l0 + l1 * l1 + l2 * l2);
    rsi = 1.0f / C_d;
  }
  s = mk(0.0f, 0.0f, 0.0f);
  w = 0.0f;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const f3 h = hcol(j) * rsi;
    w = w + A::sq_len(h) * m[j + 1];
    s = s + h;
  }
  w = w + A::sq_len(s) * m[0];
  if (w < kFeps) return 0;
  const float dl_d = (A::div(-C_d, w + (FAST ? k.alpha_d * inv_rest : k.alpha_d / rest)) * rsi) * scale;
  s = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const f3 t = gcol(j) * dl_h + hcol(j) * dl_d;
    d[j + 1] = t * m[j + 1];
    s = s + t;
  }
  d[0] = -s * m[0];
  return 1;
}

// gravity (integrator.py:40-63): v' = v + g*w*dt*(1-damp) (v' = 0 when w < eps);
// returns p = x + v'*dt.
__device__ __forceinline__ f3 predict(f3 x, f3 v, float w, f3 g, float dt, float damp1) {
  v = v + ((g * w) * dt) * damp1;
  if (w < kFeps) v = mk(0.0f, 0.0f, 0.0f);
  return x + v * dt;
}

// closestPointOnEdge (integrator.py:1521-1534).
__device__ __forceinline__ f3 closest_on_edge(f3 p, f3 ea, f3 eb) {
  const f3 d = eb - ea;
  float t;
  if (length(d) < kFeps) {
    t = 0.5f;
  } else {
    const float d2 = length(d) * length(d);
    t = dot(d, p - ea) / d2;
    if (t < 0.0f) t = 0.0f;
    else if (t > 1.0f) t = 1.0f;
  }
  return ea + d * t;
}

// triangle_closest_point_barycentric (integrator.py:1536-1582).
__device__ __forceinline__ f3 closest_bary(f3 a, f3 b, f3 c, f3 p) {
  const f3 ab = b - a, ac = c - a, ap = p - a;
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return mk(1.0f, 0.0f, 0.0f);
  const f3 bp = p - b;
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return mk(0.0f, 1.0f, 0.0f);
  const float vc = d1 * d4 - d3 * d2;
  float v = d1 / (d1 - d3);
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return mk(1.0f - v, v, 0.0f);
  const f3 cp = p - c;
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return mk(0.0f, 0.0f, 1.0f);
  const float vb = d5 * d2 - d1 * d6;
  float w = d2 / (d2 - d6);
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return mk(1.0f - w, 0.0f, w);
  const float va = d3 * d6 - d5 * d4;
  w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) return mk(0.0f, w, 1.0f - w);
  const float denom = 1.0f / (va + vb + vc);
  v = vb * denom;
  w = vc * denom;
  return mk(1.0f - v - w, v, w);
}

// checkSdfCapsuleCollision (integrator.py:1303-1355): adds the responses of
// the triangle's vertices to one capsule to r[0..2]; returns the sum of their
// lengths (0: no contact).  As in the reference, the barycentric coordinates
// of the line/plane intersection, not the point, are what closestPointOnEdge
// receives.
__device__ __forceinline__ float capsule_response(f3 tip, f3 base, float radius, f3 tri_n, f3 tri_c, f3 pA, f3 pB,
                                                  f3 pC, float ks, f3 r[3]) {
  const f3 cn = normalize(tip - base);
  const f3 off = cn * radius;
  const f3 posA = base + off, posB = tip - off;
  const float denom = dot(tri_n, cn);
  if (fabsf(denom) < kFeps) return 0.0f;
  const float t = dot(tri_n, tri_c - base) / denom;
  const f3 lpi = base + cn * t;
  const f3 refp = closest_bary(pA, pB, pC, lpi);
  const f3 sc = closest_on_edge(refp, posA, posB);
  const f3 bary = closest_bary(pA, pB, pC, sc);
  const f3 contact = (pA * bary.x + pB * bary.y) + pC * bary.z;
  const f3 sv = sc - contact;
  const float distance = length(sv);
  const float pen = radius / length(sv) - 1.0f;
  const f3 pv = mk(pen * sv.x, pen * sv.y, pen * sv.z);
  const float plen = length(pv);
  const f3 pn = normalize(pv);
  const float s = bary.x * bary.x + bary.y * bary.y + bary.z * bary.z;
  if (s < kFeps) return 0.0f;
  const f3 resp = (pn * (plen / s)) * ks;
  if (distance > radius) return 0.0f;
  const f3 ra = resp * bary.x, rb = resp * bary.y, rc = resp * bary.z;
  r[0] = r[0] + ra;
  r[1] = r[1] + rb;
  r[2] = r[2] + rc;
  return (length(ra) + length(rb)) + length(rc);
}

// Box around the points a capsule response can come from: the segment
// posA..posB of capsule_response grown by the radius (and a relative 1e-5
// plus 1e-6 absolute margin over rounding).  A triangle whose box misses it
// is farther than the radius from every point of the segment, so its
// response is zero: skipping it changes no result.
struct CapsuleBox {
  f3 lo, hi;
};

__device__ __forceinline__ CapsuleBox capsule_box(f3 tip, f3 base, float radius) {
  const f3 cn = normalize(tip - base);
  const f3 off = cn * radius;
  const f3 a = base + off, b = tip - off;
  const float m = radius * 1.00001f + 1e-6f;
  return {mk(fminf(a.x, b.x) - m, fminf(a.y, b.y) - m, fminf(a.z, b.z) - m),
          mk(fmaxf(a.x, b.x) + m, fmaxf(a.y, b.y) + m, fmaxf(a.z, b.z) + m)};
}

__device__ __forceinline__ bool tri_meets_box(const CapsuleBox& c, f3 pA, f3 pB, f3 pC) {
  const float x0 = fminf(fminf(pA.x, pB.x), pC.x), x1 = fmaxf(fmaxf(pA.x, pB.x), pC.x);
  const float y0 = fminf(fminf(pA.y, pB.y), pC.y), y1 = fmaxf(fmaxf(pA.y, pB.y), pC.y);
  const float z0 = fminf(fminf(pA.z, pB.z), pC.z), z1 = fmaxf(fmaxf(pA.z, pB.z), pC.z);
  return x1 >= c.lo.x && x0 <= c.hi.x && y1 >= c.lo.y && y0 <= c.hi.y && z1 >= c.lo.z && z0 <= c.hi.z;
}

// dragConstraintsDO (integrator.py:412-442): p -= |p-tip| * n * drag * ks / 2.
template <bool FAST>
__device__ __forceinline__ f3 grasp_pull(f3 p, f3 tip, float dr, float ks) {
  f3 dir = p - tip;
  float len = Ar<FAST>::sqrt(dot(dir, dir));
  if (len < kFeps) return p;
  if (FAST) return p - dir * ((dr * ks) * 0.5f);  // |dir| * dir/|dir| = dir
  f3 e = (((normalize(dir) * len) * dr) * ks) / 2.0f;
  return p - e;
}

}  // namespace ffsrl
