// Physics kernels: the batched frame kernel and the plain per-kernel form.
//
// Batched form (ffsrl_frame): one workgroup owns one environment for a whole
// frame.  The Jacobi deltas and constraint counts of the env's vertices are
// reduced in LDS with 64-bit integer ds_add (16 B per vertex: fixed-point dP
// and the count), so no global atomics are issued; phases (predict, project,
// apply) are separated by workgroup barriers instead of kernel boundaries, and
// the integrate of
// substep s is fused with the predict of substep s+1.  The BVH refit of the
// env runs at the end of the same workgroup.  Environments of any mix of sizes
// run in one launch; dispatching them largest-first balances the CUs.
//
// Plain form (ffsrl_predict ... ffsrl_bounds_integrate): the reference's
// launch model (integrator.py:1644-1741), one launch per kernel per env, dP
// and counts accumulated with global float / int atomics.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "constraints.hpp"
#include "refit.hpp"

namespace ffsrl {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  return fail(static_cast<int>(e), "%s: %s", what, hipGetErrorString(e));
}

// LDS budget of one workgroup on gfx950.
constexpr int kLdsBytes = 160 * 1024;
constexpr int lds_bytes_for(int nv) { return nv * 16; }
// 512 B are left for the kernels' static LDS (scheduler and pass words, the
// refit's 65 group offsets: 272-288 B today)
constexpr int kMaxFrameVertices = (kLdsBytes - 512) / 16;

// Jacobi sums in LDS are fixed point: each delta component is rounded to the
// nearest multiple of the env's quantum q = 2^dp_exp (2^-26 of the mesh
// diagonal D rounded down to a power of two) and added with an integer LDS
// atomic.  Integer atomics run at
// the LDS's full rate on gfx950 (float LDS atomics are ~30x slower), and
// integer sums do not depend on the order of the additions, so the frame is
// deterministic.  Each term is clamped to +-(2^24 - 1) q (D/8 .. D/4), so the
// sums of up to FFSRL_MAX_TERMS = 128 terms (a vertex's incident constraints
// in one pass, checked at launch) stay inside int32; the rounding error per
// term is <= q/2, 1/16 of the float32 spacing of a coordinate of size D.  A
// term beyond the range, or not finite, is clamped and flags the env
// (FFSRL_STATUS_SATURATED): the frame then no longer equals the plain form,
// which sums floats.
//
// A vertex owns 16 B of LDS, two 64-bit words updated with ds_add_u64:
//   w0 = X + Y * 2^32,  w1 = Z + count * 2^32   (two's complement, mod 2^64)
// where X, Y, Z are the int32 fixed-point sums.  Integer addition mod 2^64 is
// exact and associative, so w0 and w1 hold exactly these values whatever the
// order; since |X| < 2^31, X is the low half of w0 read as int32 and
// Y = high half - (X < 0 ? -1 : 0) (likewise Z and the count).  One 64-bit
// atomic thus carries two sums at one address.
constexpr float kFixMax = 16777215.0f;  // 2^24 - 1 quanta

struct FrameCtx {
  unsigned long long* acc;  // 2 words per vertex
  float inv_q;              // 2^-dp_exp
  float q;                  // 2^dp_exp
};

// A wavefront-uniform float held in a scalar register (its value unchanged),
// so that constants of the projection do not take vector registers.
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ FrameCtx uniform(FrameCtx c) { return FrameCtx{c.acc, uniform(c.inv_q), uniform(c.q)}; }
__device__ __forceinline__ NHConst uniform(NHConst k) {
  return NHConst{uniform(k.gamma), uniform(k.alpha_h), uniform(k.alpha_d)};
}

// A term that left the range or was not finite adds kSatMark to its count
// (counts stay below 2^16: at most FFSRL_MAX_TERMS terms of weight <= 2), so
// the apply phase sees it in the sum's count and sets the workgroup's
// saturation flag (one LDS word, cleared and read once per env frame); no
// register holds a flag across the projection loops.
constexpr uint32_t kSatMark = 1u << 16;

__device__ __forceinline__ int& sat_flag() {
  __shared__ int flag;
  return flag;
}

// The count of a decoded sum without saturation marks; raises the flag.
__device__ __forceinline__ int32_t sum_count(int32_t n) {
  if (__builtin_expect((uint32_t)n >= kSatMark, 0)) sat_flag() = 1;
  return n & (int32_t)(kSatMark - 1u);
}

__device__ __forceinline__ int32_t to_fix(float v) {
  v = rintf(v);
  v = fminf(fmaxf(v, -kFixMax), kFixMax);
  return (int32_t)v;
}
// (lo + hi * 2^32) as a 64-bit two's complement addend, lo sign-extended.
__device__ __forceinline__ unsigned long long pack2(int32_t lo, int32_t hi) {
  const uint32_t h = (uint32_t)hi + (uint32_t)(lo >> 31);
  return ((unsigned long long)h << 32) | (uint32_t)lo;
}
__device__ __forceinline__ void unpack2(unsigned long long w, int32_t* lo, int32_t* hi) {
  const int32_t l = (int32_t)(uint32_t)w;
  *lo = l;
  *hi = (int32_t)((uint32_t)(w >> 32) - (uint32_t)(l >> 31));
}
// Adds sign * d (given in quanta: the delta times c.inv_q) to the fixed-point
// sums of vertex i and n to its count.
template <int SIGN>
__device__ __forceinline__ void lds_acc(const FrameCtx& c, int i, f3 d, uint32_t n) {
  const float sx = d.x, sy = d.y, sz = d.z;
  // IEEE 754-2019 maximum propagates NaN, which fails the comparison
  const float m = __builtin_elementwise_maximum(__builtin_elementwise_maximum(fabsf(sx), fabsf(sy)), fabsf(sz));
  const uint32_t mark = m <= kFixMax ? 0u : kSatMark;
  const int32_t x = to_fix(sx), y = to_fix(sy), z = to_fix(sz);
  atomicAdd(&c.acc[2 * i], pack2(SIGN * x, SIGN * y));
  atomicAdd(&c.acc[2 * i + 1], pack2(SIGN * z, (int32_t)(n + mark)));
}

__device__ __forceinline__ float4 with_w(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// Predict + (optionally) everything of the apply phase for one vertex.
struct VertexStep {
  f3 g;
  float dt, damp1, inv_dt, ground, ks;
};

// Topology of one tet / edge of the projection loop.
struct TetIn {
  int4 q;
  float rest;
};
struct EdgeIn {
  int2 q;
  float rest;
};

// The frame reads the 16-bit packed ids (half the index bytes of int4 / int2;
// an env of the batched frame has at most 10,208 vertices, a part of the
// partitioned frame at most kPartMaxLocal).
__device__ __forceinline__ TetIn load_tet(const ffsrl_batch& b, const uint64_t* t16, int64_t ti) {
  TetIn t;
  const uint64_t w = t16[ti];
  t.q = make_int4((int)(w & 0xffffu), (int)((w >> 16) & 0xffffu), (int)((w >> 32) & 0xffffu), (int)(w >> 48));
  t.rest = b.tet_rest_volume[ti];
  return t;
}
__device__ __forceinline__ EdgeIn load_edge(const ffsrl_batch& b, const uint32_t* e16, int64_t ei) {
  EdgeIn t;
  const uint32_t w = e16[ei];
  t.q = make_int2((int)(w & 0xffffu), (int)(w >> 16));
  t.rest = b.edge_rest[ei];
  return t;
}

// Positions and inverse rest matrix of one tet, gathered ahead of use.
struct TetG {
  float4 P[4];
  float M[9];
};

// Where the projection reads positions, by mesh-local vertex id: the env's
// rows of a global array (plain, or coherent for state other workgroups
// write), or the env's copy in LDS.
template <bool COH>
struct PosGlobal {
  const float* P;
  int64_t vb;
  __device__ __forceinline__ float4 operator()(int q) const { return ldp<COH>(P, vb + q); }
};
struct PosLds {
  const float4* L;
  __device__ __forceinline__ float4 operator()(int q) const { return L[q]; }
};

template <class PS>
__device__ __forceinline__ TetG gather_tet(const ffsrl_batch& b, const PS& ps, const TetIn& t, int64_t ti, int64_t T,
                                           int nh) {
  TetG g;
  // the position gathers first: they need only the ids, and the inverse
  // rest matrix loads then overlap them
  g.P[0] = ps(t.q.x);
  g.P[1] = ps(t.q.y);
  g.P[2] = ps(t.q.z);
  g.P[3] = ps(t.q.w);
  if (nh) {
#pragma unroll
    for (int k = 0; k < 9; ++k) g.M[k] = b.tet_inv_rest[k * T + ti];
  }
  return g;
}

// Volume + Neo-Hookean deltas of one tet, summed into LDS.  fresh: the
// carried lambda starts from zero (first iteration of a substep) instead of
// being read.
template <bool CARRY, bool COH>
__device__ __forceinline__ void process_tet(const ffsrl_batch& b, const ffsrl_params& prm, const ffsrl_env_desc& e,
                                            const FrameCtx& c, const NHConst& nhk, const TetIn& tc, const TetG& g,
                                            int j, bool fresh, bool nh) {
  f3 d[4] = {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
  uint32_t n = 0;
  if (prm.vol_mode == FFSRL_PBD) {
    n += volume_pbd<true>(g.P, tc.rest, prm.vol_k, d, c.inv_q);
  } else if (prm.vol_mode == FFSRL_XPBD) {
    float lam = 0.0f;
    if (CARRY && !fresh) lam = COH ? ldfc(b.lam_t, e.lam_tbase + j) : b.lam_t[e.lam_tbase + j];
    n += volume_xpbd<true>(g.P, tc.rest, prm.vol_k, &lam, d, c.inv_q);
    if (CARRY) b.lam_t[e.lam_tbase + j] = lam;
  }
  if (nh) {
    f3 dn[4];
    if (neohookean<true>(g.P, tc.rest, g.M, nhk, dn, c.inv_q)) {
      n += 1;
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k] = d[k] + dn[k];
    }
  }
  if (n) {
    lds_acc<1>(c, tc.q.x, d[0], n);
    lds_acc<1>(c, tc.q.y, d[1], n);
    lds_acc<1>(c, tc.q.z, d[2], n);
    lds_acc<1>(c, tc.q.w, d[3], n);
  }
}

// Distance delta of one edge, summed into LDS.
template <bool CARRY, bool COH>
__device__ __forceinline__ void process_edge(const ffsrl_batch& b, const ffsrl_params& prm, const ffsrl_env_desc& e,
                                             const FrameCtx& c, const EdgeIn& ec, float4 A, float4 B, int j,
                                             bool fresh) {
  f3 d;
  int ok;
  if (prm.dist_mode == FFSRL_PBD) {
    ok = distance_pbd<true>(A, B, ec.rest, prm.dist_k, &d, c.inv_q);
  } else {
    float lam = 0.0f;
    if (CARRY && !fresh) lam = COH ? ldfc(b.lam_e, e.lam_ebase + j) : b.lam_e[e.lam_ebase + j];
    ok = distance_xpbd<true>(A, B, ec.rest, prm.dist_k, &lam, &d, c.inv_q);
    // an early-out leaves lambda unchanged: a fresh one is stored as zero
    if (CARRY && (ok || fresh)) b.lam_e[e.lam_ebase + j] = lam;
  }
  if (ok) {
    lds_acc<-1>(c, ec.q.x, d * A.w, 1u);
    lds_acc<1>(c, ec.q.y, d * B.w, 1u);
  }
}

// The constraints of a projection pass: packed ids (the batch's mesh-local
// ids, or the partitioned frame's part-local ids) and the first tet and edge
// of the range within the env's mesh (item j is tet t0 + j and edge e0 + j;
// the carried lambdas are indexed the same way).
struct Items {
  const uint64_t* t16;
  const uint32_t* e16;
  int t0, e0;
};

// Work of one stride for one thread: tet j and edge j of the env.
struct Slot {
  TetIn t;
  TetG g;
  EdgeIn e;
  float4 A, B;
};

// The 1024-thread projection loop for a fixed constraint set (HT: tets,
// HE: edges, NH: Neo-Hookean): indices are clamped into range instead of
// guarded, so every load of the loop body is unconditional and no branch
// lies between a load and its use (a branch that issues loads on one path
// only makes the compiler wait for every outstanding load at the join).
// The ids are loaded one stride ahead; an item's edge gathers, tet gathers
// and inverse rest matrix are issued together, then the edge and the tet
// are computed (the edge's registers are free again before the tet's peak).
template <bool HT, bool HE, bool NH, bool CARRY, bool COH, class PS>
__device__ __forceinline__ void project_wide(const ffsrl_batch& b, const ffsrl_params& prm, const ffsrl_env_desc& e,
                                             const FrameCtx& c, const NHConst& nhk, const PS& ps, const Items& it,
                                             int64_t T, int nt, int ne, int jb, int je, bool fresh) {
  constexpr int BLOCK = 1024;
  const int nmax = je;
  const int j0 = jb + threadIdx.x;
  auto ct = [&](int jj) { return jj < nt ? jj : nt - 1; };
  auto ce = [&](int jj) { return jj < ne ? jj : ne - 1; };
  const int tb = e.tbase + it.t0, eb = e.ebase + it.e0;  // 32-bit like the env table
  auto ids = [&](Slot& s, int jj) {
    if (HT) s.t = load_tet(b, it.t16, tb + ct(jj));
    if (HE) s.e = load_edge(b, it.e16, eb + ce(jj));
  };
  auto gr = [&](const Slot& s, int jj) {
    float4 A = make_float4(0.0f, 0.0f, 0.0f, 0.0f), B = A;
    TetG g;
    if (HE) {
      A = ps(s.e.q.x);
      B = ps(s.e.q.y);
    }
    if (HT) g = gather_tet(b, ps, s.t, tb + ct(jj), T, NH);
    if (HE && jj < ne) process_edge<CARRY, COH>(b, prm, e, c, s.e, A, B, it.e0 + jj, fresh);
    if (HT && jj < nt) process_tet<CARRY, COH>(b, prm, e, c, nhk, s.t, g, it.t0 + jj, fresh, NH);
  };
  Slot s0, s1;
  ids(s0, j0);
  for (int j = j0; j < nmax; j += 2 * BLOCK) {
    ids(s1, j + BLOCK);
    gr(s0, j);
    if (j + BLOCK >= nmax) break;
    ids(s0, j + 2 * BLOCK);
    gr(s1, j + BLOCK);
  }
}

// One projection pass over the env's tets and edges.  Thread j handles tet j
// and edge j of each stride of BLOCK.  The pass is software-pipelined over
// rotating slots (the loop is unrolled over the slots, so "next" and
// "current" are different registers and no load result is copied before it
// is consumed):
//   BLOCK <= 512 (256 VGPRs per lane): three slots; indices are loaded two
//     strides ahead and positions / inverse rest matrices one stride ahead;
//   BLOCK == 1024 (128 VGPRs per lane): two slots; indices are loaded one
//     stride ahead and positions at use.
//
// Items jb .. je-1 (tets below nt, edges below ne); the batched frame passes
// the whole range, a shared env one chunk of it.
template <int BLOCK, bool CARRY, bool COH, class PS>
__device__ __forceinline__ void project(const ffsrl_batch& b, const ffsrl_params& prm, const ffsrl_env_desc& e,
                                        const FrameCtx& c, const NHConst& nhk, const PS& ps, const Items& it,
                                        int64_t T, int nt, int ne, int jb, int je, bool fresh) {
  const int nmax = je;
  const int j0 = jb + threadIdx.x;
  const int tb = e.tbase + it.t0, eb = e.ebase + it.e0;  // 32-bit like the env table
  auto ids = [&](Slot& s, int jj) {
    if (jj < nt) s.t = load_tet(b, it.t16, tb + jj);
    if (jj < ne) s.e = load_edge(b, it.e16, eb + jj);
  };
  auto gat = [&](Slot& s, int jj) {
    if (jj < nt) s.g = gather_tet(b, ps, s.t, tb + jj, T, prm.nh_on);
    if (jj < ne) {
      s.A = ps(s.e.q.x);
      s.B = ps(s.e.q.y);
    }
  };
  auto run = [&](const Slot& s, int jj) {
    if (jj < nt) process_tet<CARRY, COH>(b, prm, e, c, nhk, s.t, s.g, it.t0 + jj, fresh, prm.nh_on != 0);
    if (jj < ne) process_edge<CARRY, COH>(b, prm, e, c, s.e, s.A, s.B, it.e0 + jj, fresh);
  };
  if constexpr (BLOCK <= 512) {
    Slot s0, s1, s2;
    ids(s0, j0);
    ids(s1, j0 + BLOCK);
    gat(s0, j0);
    for (int j = j0; j < nmax; j += 3 * BLOCK) {
      ids(s2, j + 2 * BLOCK);
      gat(s1, j + BLOCK);
      run(s0, j);
      if (j + BLOCK >= nmax) break;
      ids(s0, j + 3 * BLOCK);
      gat(s2, j + 2 * BLOCK);
      run(s1, j + BLOCK);
      if (j + 2 * BLOCK >= nmax) break;
      ids(s1, j + 4 * BLOCK);
      gat(s0, j + 3 * BLOCK);
      run(s2, j + 2 * BLOCK);
    }
  } else {
    // 128 VGPRs per lane: the straight-line loop of the constraint set
    const bool nh = prm.nh_on != 0;
    if (nt > 0 && ne > 0) {
      if (nh)
        project_wide<true, true, true, CARRY, COH>(b, prm, e, c, nhk, ps, it, T, nt, ne, jb, je, fresh);
      else
        project_wide<true, true, false, CARRY, COH>(b, prm, e, c, nhk, ps, it, T, nt, ne, jb, je, fresh);
    } else if (nt > 0) {
      if (nh)
        project_wide<true, false, true, CARRY, COH>(b, prm, e, c, nhk, ps, it, T, nt, ne, jb, je, fresh);
      else
        project_wide<true, false, false, CARRY, COH>(b, prm, e, c, nhk, ps, it, T, nt, ne, jb, je, fresh);
    } else if (ne > 0) {
      project_wide<false, true, false, CARRY, COH>(b, prm, e, c, nhk, ps, it, T, nt, ne, jb, je, fresh);
    }
  }
}

// Surface triangles of an env: mesh-local (a, b, c) from the batch's int4
// triangles, or part-local ids packed 3 x 16 bits (the partitioned frame).
struct TrisGlobal {
  const int32_t* tris;
  int64_t base;
  __device__ __forceinline__ int3 operator()(int f) const {
    const int4 t = ldi4(tris, base + f);
    return make_int3(t.x, t.y, t.z);
  }
};
struct TrisPacked {
  const uint64_t* t16;
  int64_t base;
  __device__ __forceinline__ int3 operator()(int f) const {
    const uint64_t w = t16[base + f];
    return make_int3((int)(w & 0xffffu), (int)((w >> 16) & 0xffffu), (int)((w >> 32) & 0xffffu));
  }
};

// collideLaparoscopeDO (integrator.py:981-1050) for one env: every surface
// triangle against capsules 0..2 of the env's instrument, responses read from
// the positions at the start of the pass and summed (negated) into the LDS
// accumulators for the vertices with inverse mass > 0; *contacts counts the
// triangles with a response.  ps(q) is the position (w: inverse mass) of
// vertex q and ts(f) triangle f's vertex ids in ps's and the accumulators'
// numbering; triangles f0 .. f1-1 (the whole surface, one chunk of a shared
// env, or a part's triangles).
template <class PS, class TS>
__device__ __forceinline__ void collide_tris(const ffsrl_batch& b, const ffsrl_params& prm, int env,
                                             const FrameCtx& c, const PS& ps, const TS& ts, int* contacts, int block,
                                             int f0, int f1) {
  CapsuleBox box[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int ci = env * 4 + i;
    box[i] = capsule_box(mk(b.tips[3 * ci], b.tips[3 * ci + 1], b.tips[3 * ci + 2]),
                         mk(b.caps_base[3 * ci], b.caps_base[3 * ci + 1], b.caps_base[3 * ci + 2]),
                         b.caps_radius[ci]);
  }
  for (int f = f0 + (int)threadIdx.x; f < f1; f += block) {
    const int3 t = ts(f);
    const float4 A = ps(t.x), B = ps(t.y), Cc = ps(t.z);
    const f3 pA = xyz(A), pB = xyz(B), pC = xyz(Cc);
    const bool near0 = tri_meets_box(box[0], pA, pB, pC), near1 = tri_meets_box(box[1], pA, pB, pC),
               near2 = tri_meets_box(box[2], pA, pB, pC);
    if (!(near0 || near1 || near2)) continue;  // every response would be zero
    const f3 centre = ((pA + pB) + pC) * 0.3333f;
    const f3 n = normalize(cross(pB - pA, pC - pA));
    f3 r[3] = {mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f)};
    float len = 0.0f;
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
      if (!(i == 0 ? near0 : i == 1 ? near1 : near2)) continue;  // zero response
      const int ci = env * 4 + i;
      const f3 tip = mk(b.tips[3 * ci], b.tips[3 * ci + 1], b.tips[3 * ci + 2]);
      const f3 base = mk(b.caps_base[3 * ci], b.caps_base[3 * ci + 1], b.caps_base[3 * ci + 2]);
      len = len + capsule_response(tip, base, b.caps_radius[ci], n, centre, pA, pB, pC, prm.contact_k, r);
    }
    if (len > 0.0f) {
      atomicAdd(contacts, 1);
      if (A.w > 0.0f) lds_acc<-1>(c, t.x, r[0] * c.inv_q, 0u);
      if (B.w > 0.0f) lds_acc<-1>(c, t.y, r[1] * c.inv_q, 0u);
      if (Cc.w > 0.0f) lds_acc<-1>(c, t.z, r[2] * c.inv_q, 0u);
    }
  }
}

// The collision pass of env `env` over triangles f0 .. f1-1, positions from
// the model's p array.
template <bool COH>
__device__ __forceinline__ void collide_env(const ffsrl_batch& b, const ffsrl_params& prm, const ffsrl_env_desc& e,
                                            int env, const FrameCtx& c, const float* P, int64_t vb, int* contacts,
                                            int block, int f0, int f1) {
  collide_tris(b, prm, env, c, PosGlobal<COH>{P, vb}, TrisGlobal{b.tris, (int64_t)e.fbase}, contacts, block, f0, f1);
}

// One frame of env `env` by the calling workgroup (COLLIDE: with the
// instrument collision pass).
template <int BLOCK, bool CARRY, bool COLLIDE>
__device__ __forceinline__ void frame_env(const ffsrl_batch& b, const ffsrl_params& prm, int env, int refit,
                                          unsigned long long* lds) {
  const ffsrl_env_desc e = b.envs[env];
  const int nv = e.nv;
  const FrameCtx c = uniform(FrameCtx{lds, ldexpf(1.0f, -e.dp_exp), ldexpf(1.0f, e.dp_exp)});
  const int tid = threadIdx.x;
  if (tid == 0) sat_flag() = 0;  // read after the frame's barriers
  const int64_t vb = e.vbase;
  float* X = b.x;
  float* P = b.p;
  float* Vv = b.v;

  VertexStep vs;
  vs.g = mk(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
  vs.dt = prm.dt;
  vs.damp1 = 1.0f - prm.damping;
  vs.inv_dt = 1.0f / prm.dt;
  vs.ground = prm.ground;
  vs.ks = prm.drag_ks;
  const NHConst nhk = uniform(nh_const(prm.nh_vol_c, prm.nh_dev_c, prm.dt));
  const f3 tip = mk(b.tips[(env * 4 + 2) * 3 + 0], b.tips[(env * 4 + 2) * 3 + 1], b.tips[(env * 4 + 2) * 3 + 2]);
  const int64_t T = b.n_tet_total;

  for (int i = tid; i < 2 * nv; i += BLOCK) c.acc[i] = 0ull;
  // predict of substep 0; each thread notes which of its vertices (i = tid +
  // k * BLOCK, k < 64 since nv <= 10,208) are grasped, so the apply phases
  // read the grasp weight of those only
  uint64_t grasped = 0;
  for (int i = tid, k = 0; i < nv; i += BLOCK, ++k) {
    float4 x = ld4(X, vb + i), v = ld4(Vv, vb + i);
    st4(P, vb + i, with_w(predict(xyz(x), xyz(v), x.w, vs.g, vs.dt, vs.damp1), x.w));
    if (b.drag[vb + i] != 0.0f) grasped |= 1ull << k;
  }
  __syncthreads();

  const bool reset_lambda = !prm.lambda_legacy || prm.iterations > 1;
  constexpr bool collide = COLLIDE;
  __shared__ int contacts;  // surface triangles in contact during the frame
  if (tid == 0) contacts = 0;
  const bool do_tets = prm.vol_mode != FFSRL_OFF || prm.nh_on;
  const bool do_edges = prm.dist_mode != FFSRL_OFF;
  for (int s = 0; s < prm.substeps; ++s) {
    if (CARRY && reset_lambda) {
      // Each thread zeroes the lambdas of exactly the constraints it projects.
      if (prm.vol_mode == FFSRL_XPBD)
        for (int t = tid; t < e.nt; t += BLOCK) b.lam_t[e.lam_tbase + t] = 0.0f;
      if (prm.dist_mode == FFSRL_XPBD)
        for (int k = tid; k < e.ne; k += BLOCK) b.lam_e[e.lam_ebase + k] = 0.0f;
    }
    for (int it = 0; it < prm.iterations; ++it) {
      // ---- projection: tets (volume + Neo-Hookean) and edges (distance) ----
      {
        const int nt = do_tets ? e.nt : 0, ne = do_edges ? e.ne : 0;
        project<BLOCK, CARRY, false>(b, prm, e, c, nhk, PosGlobal<false>{P, vb}, Items{b.tets16, b.edges16, 0, 0},
                                     T, nt, ne, 0, nt > ne ? nt : ne, false);
      }
      __syncthreads();
      // ---- apply (+ grasp; + ground, integrate and next predict) ----
      const bool last_it = (it == prm.iterations - 1);
      const bool next_sub = (s + 1 < prm.substeps);
      // ground clamp, integrate and the next substep's predict of vertex i
      auto finish = [&](int i, f3 p, float w) {
        if (p.y < vs.ground) p.y = vs.ground;
        const float4 x4 = ld4(X, vb + i);
        const f3 v = (p - xyz(x4)) * vs.inv_dt;
        st4(X, vb + i, with_w(p, x4.w));
        // v feeds the next substep's predict here; only the frame's last is state
        if (next_sub)
          p = predict(p, v, x4.w, vs.g, vs.dt, vs.damp1);
        else
          st4(Vv, vb + i, with_w(v, 0.0f));
        st4(P, vb + i, with_w(p, w));
      };
      const bool fuse = last_it && !collide;  // with collision, the integrate follows it
      for (int i = tid, k = 0; i < nv; i += BLOCK, ++k) {
        const unsigned long long w0 = c.acc[2 * i], w1 = c.acc[2 * i + 1];
        c.acc[2 * i] = 0ull;
        c.acc[2 * i + 1] = 0ull;
        int32_t fx, fy, fz, n;
        unpack2(w0, &fx, &fy);
        unpack2(w1, &fz, &n);
        n = sum_count(n);
        float4 p4 = ld4(P, vb + i);
        f3 p = xyz(p4);
        if (n > 0) {
          const float s = c.q * __builtin_amdgcn_rcpf((float)n);
          p = p + mk((float)fx, (float)fy, (float)fz) * s;
        }
        if ((grasped >> k) & 1ull) p = grasp_pull<true>(p, tip, b.drag[vb + i], vs.ks);
        if (!fuse) {
          st4(P, vb + i, with_w(p, p4.w));
          continue;
        }
        finish(i, p, p4.w);
      }
      __syncthreads();
      if (collide) {
        // ---- instrument collision (collideLaparoscopeDO), summed in LDS ----
        collide_env<false>(b, prm, e, env, c, P, vb, &contacts, BLOCK, 0, e.nf);
        __syncthreads();
        for (int i = tid; i < nv; i += BLOCK) {
          const unsigned long long w0 = c.acc[2 * i], w1 = c.acc[2 * i + 1];
          c.acc[2 * i] = 0ull;
          c.acc[2 * i + 1] = 0ull;
          if (!last_it && (w0 | w1) == 0ull) continue;
          int32_t fx, fy, fz, n;
          unpack2(w0, &fx, &fy);
          unpack2(w1, &fz, &n);
          (void)sum_count(n);  // collision terms carry no count, only saturation marks
          const float4 p4 = ld4(P, vb + i);
          const f3 p = xyz(p4) + mk((float)fx, (float)fy, (float)fz) * c.q;
          if (last_it)
            finish(i, p, p4.w);
          else
            st4(P, vb + i, with_w(p, p4.w));
        }
        __syncthreads();
      }
    }
  }
  if (collide && tid == 0) b.in_collision[env] += contacts;  // counts accumulate until the caller clears them
  __syncthreads();
  if (tid == 0 && sat_flag()) atomicOr(&b.status[env], (int32_t)FFSRL_STATUS_SATURATED);
  if (refit) refit_env<BLOCK>(b, e);
}

// Envs are taken from a work counter (sched[0]) in `order`, so a workgroup
// that finishes an env starts the next one at once: the launch has one
// workgroup per CU and balances like greedy list scheduling over the
// largest-first order, with no workgroup dispatch between envs.  The last
// workgroup to finish (sched[1]) resets both counters for the next launch.
// With sched == NULL each workgroup takes env order[blockIdx.x].
template <int BLOCK, bool CARRY, bool COLLIDE>
__global__ __launch_bounds__(BLOCK) void k_frame(ffsrl_batch b, ffsrl_params prm, const int32_t* order, int n,
                                                 int refit, uint32_t* sched) {
  extern __shared__ unsigned long long lds[];
  if (!sched) {
    frame_env<BLOCK, CARRY, COLLIDE>(b, prm, order ? order[blockIdx.x] : (int)blockIdx.x, refit, lds);
    return;
  }
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = (int)atomicAdd(&sched[0], 1u);
    __syncthreads();
    const int k = next;
    __syncthreads();
    if (k >= n) break;  // uniform: every thread read the same k
    frame_env<BLOCK, CARRY, COLLIDE>(b, prm, order ? order[k] : k, refit, lds);
    __syncthreads();  // the env's LDS and phases are done before the next env
  }
  if (threadIdx.x == 0 && atomicAdd(&sched[1], 1u) == gridDim.x - 1) {
    atomicExch(&sched[0], 0u);
    atomicExch(&sched[1], 0u);
  }
}

template <int BLOCK>
static int launch_frame(const ffsrl_batch& b, const ffsrl_params& prm, const int32_t* order, int n, int refit,
                        bool carry, int lds, uint32_t* sched, hipStream_t st) {
  auto kern = prm.collide ? (carry ? k_frame<BLOCK, true, true> : k_frame<BLOCK, false, true>)
                          : (carry ? k_frame<BLOCK, true, false> : k_frame<BLOCK, false, false>);
  FFSRL_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  int grid = n;
  if (sched) {
    // one resident workgroup per CU at most (the frame uses up to all LDS)
    int dev = 0, cus = 0, per_cu = 0;
    FFSRL_HIP(hipGetDevice(&dev));
    FFSRL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    FFSRL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), BLOCK, lds));
    grid = std::min(n, std::max(1, cus * std::max(1, per_cu)));
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), lds, st, b, prm, order, n, refit, sched);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Shared-env frame (ffsrl_frame_shared): the workgroups of one XCD share the
// frames of that XCD's environments (list slot el belongs to XCD el % nxcd,
// XCDs numbered densely from the hardware ids each workgroup reads).
//
// The frame of an env is a fixed sequence of passes (predict; then per
// substep and iteration: projection, apply [, collision, collision apply]),
// each cut into chunks numbered in pass order.  Workgroups take chunks from
// the env's counter ctr[0]; a chunk of a pass starts once every chunk of the
// passes before it is done (ctr[1] >= the pass's first chunk).  Projection
// and collision chunks sum into the workgroup's LDS accumulators; the
// workgroup adds its sums into the env's global sums (split_acc, the same
// 64-bit fixed-point words as the LDS form, with integer atomics that the
// XCD's L2 performs) when it leaves the pass, and only then counts those
// chunks done.  An apply chunk reads and clears the global sums of its
// vertices.  Integer sums do not depend on the order of the additions, so the
// result equals the one-workgroup frame bit for bit.
//
// Progress: a workgroup waits only for the passes before its current chunk,
// and only after it has counted every chunk it has finished; only running
// workgroups hold chunks, so the lowest chunk not yet counted always belongs
// to a workgroup that can run it, and every pass completes whatever number
// of workgroups is resident.  Waits are bounded anyway (kShareSpinLimit): a
// workgroup that gives up sets the env's give-up flag ctr[2], which makes
// the env's other waits give up too; the last workgroup out then reports
// the frame (FFSRL_STATUS_INCOMPLETE), clears its partial global sums and
// resets its counters.
//
// Coherence is an assumption about the hardware, not a guarantee of the
// HIP memory model; it is what a probe of the MI355X showed, and the
// bit-equality tests against the one-workgroup frame are its standing check.
//
// Coherence: all workgroups of an env run on one XCD and share its L2.
// Stores and atomics have reached the L2 when `s_waitcnt vmcnt(0)` returns
// (wait_vmem, before a chunk is counted done); state written during the
// kernel (positions, velocities, carried lambdas, the global sums) is read
// with nontemporal loads, which skip the CU's L1.
constexpr int kShareItems = 2048;        // tets / edges (by index) per projection chunk
constexpr int kShareVerts = 2048;        // vertices per apply chunk
constexpr int kShareTris = 2048;         // surface triangles per collision chunk
constexpr int kShareSpinLimit = 1 << 22; // polls of a pass wait before the frame gives up

enum ShareKind { kPredictPass, kProjectPass, kApplyPass, kCollidePass, kCollideApplyPass };

struct SharePlan {
  int cv, cc, cf, unit, total;
};

__host__ __device__ inline SharePlan share_plan(const ffsrl_env_desc& e, const ffsrl_params& prm, bool collide) {
  const bool do_tets = prm.vol_mode != FFSRL_OFF || prm.nh_on, do_edges = prm.dist_mode != FFSRL_OFF;
  const int nt = do_tets ? e.nt : 0, ne = do_edges ? e.ne : 0, nmax = nt > ne ? nt : ne;
  SharePlan p;
  p.cv = (e.nv + kShareVerts - 1) / kShareVerts;
  if (p.cv < 1) p.cv = 1;
  p.cc = (nmax + kShareItems - 1) / kShareItems;
  if (p.cc < 1) p.cc = 1;
  p.cf = 0;
  if (collide) {
    p.cf = (e.nf + kShareTris - 1) / kShareTris;
    if (p.cf < 1) p.cf = 1;
  }
  p.unit = p.cc + p.cv + (collide ? p.cf + p.cv : 0);
  p.total = p.cv + prm.substeps * prm.iterations * p.unit;
  return p;
}

struct ShareChunk {
  int kind, base, local, s, it;  // pass kind, first chunk of the pass, index in the pass, substep, iteration
};

__device__ __forceinline__ ShareChunk share_chunk(const SharePlan& p, int c, int iterations) {
  ShareChunk k;
  if (c < p.cv) {
    k.kind = kPredictPass;
    k.base = 0;
    k.local = c;
    k.s = 0;
    k.it = 0;
    return k;
  }
  const int c1 = c - p.cv, u = c1 / p.unit, o = c1 - u * p.unit, ub = p.cv + u * p.unit;
  k.s = u / iterations;
  k.it = u - k.s * iterations;
  if (o < p.cc) {
    k.kind = kProjectPass;
    k.base = ub;
    k.local = o;
  } else if (o < p.cc + p.cv) {
    k.kind = kApplyPass;
    k.base = ub + p.cc;
    k.local = o - p.cc;
  } else if (o < p.cc + p.cv + p.cf) {
    k.kind = kCollidePass;
    k.base = ub + p.cc + p.cv;
    k.local = o - p.cc - p.cv;
  } else {
    k.kind = kCollideApplyPass;
    k.base = ub + p.cc + p.cv + p.cf;
    k.local = o - p.cc - p.cv - p.cf;
  }
  return k;
}

// Where the shared frame keeps an env's global sums (2 words per vertex)
// and chunk counters (taken, done, give-up flag, exits): indexed by env and
// vertex base (the batch's split_acc / split_ctr), or by list slot el, with
// slot_nv vertices per slot (a workspace of the library's own).  fold: the
// folded form's workspace (the batch's fold_ws, by vertex base; NULL: every
// env takes the unfolded form).
struct ShareWs {
  unsigned long long* acc;
  uint32_t* ctr;
  int by_slot;
  int slot_nv;
  unsigned long long* fold;
  int64_t fold_stride;
  __device__ unsigned long long* acc_of(int el, const ffsrl_env_desc& e) const {
    return acc + 2 * (by_slot ? (int64_t)el * slot_nv : (int64_t)e.vbase);
  }
  __device__ uint32_t* ctr_of(int el, int env) const { return ctr + 4 * (int64_t)(by_slot ? el : env); }
  // plane k of the folded form's workspace, at the env's first vertex
  __device__ unsigned long long* fold_plane(int k, const ffsrl_env_desc& e) const {
    return fold + 2 * ((int64_t)k * fold_stride + e.vbase);
  }
};

// Largest env of the folded form: its positions and sums share the LDS
// (32 B per vertex).
constexpr int kFoldMaxVertices = (kLdsBytes - 512) / 32;

// Polls ctr[1] until it reaches target; false when the wait is given up
// (limit reached, or another workgroup of the env gave up: ctr[2] set).
// The give-up flag is read with the first poll and every 256th, both loads
// in flight together; a pass found complete is never given up.
__device__ __forceinline__ bool share_wait(uint32_t* ctr, uint32_t target) {
  for (int k = 0; k < kShareSpinLimit; ++k) {
    const uint32_t done = atomicAdd(&ctr[1], 0u);
    const uint32_t gave_up = (k & 255) == 0 ? atomicAdd(&ctr[2], 0u) : 0u;
    if (done >= target) return true;
    if (gave_up) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return false;
}

// The env's share of one workgroup: chunks until the env's counter runs out.
template <int BLOCK, bool CARRY, bool COLLIDE>
__device__ void share_env(const ffsrl_batch& b, const ffsrl_params& prm, const ShareWs& ws, int el, int env,
                          int refit, unsigned long long* lds) {
  __shared__ int s_chunk, s_ok, s_last, s_contacts;
  const ffsrl_env_desc e = b.envs[env];
  const SharePlan pl = share_plan(e, prm, COLLIDE);
  uint32_t* ctr = ws.ctr_of(el, env);
  unsigned long long* gacc = ws.acc_of(el, e);
  const int nv = e.nv, tid = threadIdx.x;
  const int64_t vb = e.vbase;
  const FrameCtx c = uniform(FrameCtx{lds, ldexpf(1.0f, -e.dp_exp), ldexpf(1.0f, e.dp_exp)});
  float* X = b.x;
  float* P = b.p;
  float* Vv = b.v;
  VertexStep vs;
  vs.g = mk(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
  vs.dt = prm.dt;
  vs.damp1 = 1.0f - prm.damping;
  vs.inv_dt = 1.0f / prm.dt;
  vs.ground = prm.ground;
  vs.ks = prm.drag_ks;
  const NHConst nhk = uniform(nh_const(prm.nh_vol_c, prm.nh_dev_c, prm.dt));
  const f3 tip = mk(b.tips[(env * 4 + 2) * 3 + 0], b.tips[(env * 4 + 2) * 3 + 1], b.tips[(env * 4 + 2) * 3 + 2]);
  const int64_t T = b.n_tet_total;
  const bool reset_lambda = !prm.lambda_legacy || prm.iterations > 1;
  const int nt = (prm.vol_mode != FFSRL_OFF || prm.nh_on) ? e.nt : 0;
  const int ne = prm.dist_mode != FFSRL_OFF ? e.ne : 0;
  const int nmax = nt > ne ? nt : ne;

  // ground clamp, integrate and the next substep's predict of vertex i (as
  // the one-workgroup frame)
  auto finish = [&](int i, f3 p, float w, bool next_sub) {
    if (p.y < vs.ground) p.y = vs.ground;
    const float4 x4 = ld4c(X, vb + i);
    const f3 v = (p - xyz(x4)) * vs.inv_dt;
    st4(X, vb + i, with_w(p, x4.w));
    if (next_sub)
      p = predict(p, v, x4.w, vs.g, vs.dt, vs.damp1);
    else
      st4(Vv, vb + i, with_w(v, 0.0f));
    st4(P, vb + i, with_w(p, w));
  };

  for (int i = tid; i < 2 * nv; i += BLOCK) lds[i] = 0ull;
  if (tid == 0) {
    s_contacts = 0;
    sat_flag() = 0;
  }
  int pend_base = -1, pend = 0;
  for (;;) {
    __syncthreads();  // LDS and the shared words of the previous chunk are done with
    if (tid == 0) s_chunk = (int)atomicAdd(&ctr[0], 1u);
    __syncthreads();
    const int ci = s_chunk;
    ShareChunk ch = {0, 0, 0, 0, 0};
    if (ci < pl.total) ch = share_chunk(pl, ci, prm.iterations);
    if (pend && (ci >= pl.total || ch.base != pend_base)) {
      // leave the pass: add the LDS sums to the env's sums, then count the chunks
      for (int i = tid; i < nv; i += BLOCK) {
        const unsigned long long w0 = lds[2 * i], w1 = lds[2 * i + 1];
        if (w0) {
          atomicAdd(&gacc[2 * i], w0);
          lds[2 * i] = 0ull;
        }
        if (w1) {
          atomicAdd(&gacc[2 * i + 1], w1);
          lds[2 * i + 1] = 0ull;
        }
      }
      wait_vmem();
      __syncthreads();
      if (tid == 0) {
        if (s_contacts) atomicAdd(&b.in_collision[env], s_contacts);
        s_contacts = 0;
        atomicAdd(&ctr[1], (uint32_t)pend);
      }
      pend = 0;
    }
    if (ci >= pl.total) break;
    if (tid == 0) s_ok = ch.base == 0 || share_wait(ctr, (uint32_t)ch.base);
    __syncthreads();
    if (!s_ok) {
      if (tid == 0) atomicOr(&ctr[2], 1u);
      break;
    }
    const bool last_it = ch.it == prm.iterations - 1, next_sub = ch.s + 1 < prm.substeps;
    if (ch.kind == kProjectPass || ch.kind == kCollidePass) {
      if (ch.kind == kProjectPass) {
        const int jb = ch.local * kShareItems, je = jb + kShareItems < nmax ? jb + kShareItems : nmax;
        project<BLOCK, CARRY, true>(b, prm, e, c, nhk, PosGlobal<true>{P, vb}, Items{b.tets16, b.edges16, 0, 0}, T,
                                    nt, ne, jb, je, reset_lambda && ch.it == 0);
      } else {
        const int f0 = ch.local * kShareTris, f1 = f0 + kShareTris < e.nf ? f0 + kShareTris : e.nf;
        collide_env<true>(b, prm, e, env, c, P, vb, &s_contacts, BLOCK, f0, f1);
      }
      pend_base = ch.base;
      pend += 1;
      continue;
    }
    const int v0 = ch.local * kShareVerts, v1 = v0 + kShareVerts < nv ? v0 + kShareVerts : nv;
    if (ch.kind == kPredictPass) {
      for (int i = v0 + tid; i < v1; i += BLOCK) {
        const float4 x = ld4c(X, vb + i), v = ld4c(Vv, vb + i);
        st4(P, vb + i, with_w(predict(xyz(x), xyz(v), x.w, vs.g, vs.dt, vs.damp1), x.w));
      }
    } else if (ch.kind == kApplyPass) {
      const bool fuse = last_it && !COLLIDE;
      for (int i = v0 + tid; i < v1; i += BLOCK) {
        const unsigned long long w0 = ldu64c(gacc, 2 * i), w1 = ldu64c(gacc, 2 * i + 1);
        if (w0) gacc[2 * i] = 0ull;
        if (w1) gacc[2 * i + 1] = 0ull;
        int32_t fx, fy, fz, n;
        unpack2(w0, &fx, &fy);
        unpack2(w1, &fz, &n);
        n = sum_count(n);
        const float4 p4 = ld4c(P, vb + i);
        f3 p = xyz(p4);
        if (n > 0) {
          const float sc = c.q * __builtin_amdgcn_rcpf((float)n);
          p = p + mk((float)fx, (float)fy, (float)fz) * sc;
        }
        const float dr = b.drag[vb + i];
        if (dr != 0.0f) p = grasp_pull<true>(p, tip, dr, vs.ks);
        if (fuse)
          finish(i, p, p4.w, next_sub);
        else
          st4(P, vb + i, with_w(p, p4.w));
      }
    } else {  // collision apply
      for (int i = v0 + tid; i < v1; i += BLOCK) {
        const unsigned long long w0 = ldu64c(gacc, 2 * i), w1 = ldu64c(gacc, 2 * i + 1);
        if (w0) gacc[2 * i] = 0ull;
        if (w1) gacc[2 * i + 1] = 0ull;
        if (!last_it && (w0 | w1) == 0ull) continue;
        int32_t fx, fy, fz, n;
        unpack2(w0, &fx, &fy);
        unpack2(w1, &fz, &n);
        (void)sum_count(n);  // collision terms carry no count, only saturation marks
        const float4 p4 = ld4c(P, vb + i);
        const f3 p = xyz(p4) + mk((float)fx, (float)fy, (float)fz) * c.q;
        if (last_it)
          finish(i, p, p4.w, next_sub);
        else
          st4(P, vb + i, with_w(p, p4.w));
      }
    }
    wait_vmem();
    __syncthreads();
    if (tid == 0) s_last = atomicAdd(&ctr[1], 1u) + 1u == (uint32_t)pl.total;
    __syncthreads();
    if (refit && s_last) refit_env<BLOCK, true>(b, e);  // every chunk of the frame is done
  }
  __syncthreads();
  if (tid == 0 && sat_flag()) atomicOr(&b.status[env], (int32_t)FFSRL_STATUS_SATURATED);
}


// ---------------------------------------------------------------------------
// Folded shared frame (an env of at most kFoldMaxVertices vertices, without
// the collision pass).  The unfolded frame has two passes per unit (substep
// x iteration): the projection, then an apply pass that publishes the
// positions the next projection reads, each ended by a hand-off.  Here the
// apply is folded into the next pass: pass u holds the projection chunks of
// unit u and vertex chunks that publish unit u's positions; a workgroup that
// projects in pass u first computes the env's positions of unit u into LDS
// from the published positions of unit u - 1 and that unit's global sums
// (the apply and grasp and, at a substep boundary, the ground clamp,
// integrate and predict, expression by expression as the one-workgroup
// frame), then gathers from LDS.  One hand-off per unit.  A final pass of
// vertex chunks does the last apply and integrate and writes the model's x,
// v and p.
//
// Workspace planes (fold_stride vertices x 16 B, by vertex base):
//   0-2  the global sums of unit u in plane u % 3;
//   3-4  the published positions of unit u in plane 3 + u % 2;
//   5-6  the positions at the start of substep s >= 1 in plane 5 + s % 2
//        (substep 0 starts from the model's x).
// Every reader of a plane runs in a later pass than its writers, and no
// plane is rewritten before every pass that reads it has completed: pass u's
// vertex chunks clear the sums of unit u - 2 (last read in pass u - 1), and
// the final pass clears the last two, so every frame leaves the sums at
// zero.  Sums and published state are read with coherent loads.

// Positions of unit u from unit u - 1: the apply of its sums (+ grasp) and,
// at a substep boundary, the ground clamp, integrate and next predict.
// *xs gets the new substep's start position (boundary only).
__device__ __forceinline__ float4 fold_advance(const FrameCtx& c, const VertexStep& vs, const f3& tip, float4 p4,
                                               unsigned long long w0, unsigned long long w1, float dr, bool boundary,
                                               float4 x4, float4* xs) {
  int32_t fx, fy, fz, n;
  unpack2(w0, &fx, &fy);
  unpack2(w1, &fz, &n);
  n = sum_count(n);
  f3 p = xyz(p4);
  if (n > 0) {
    const float s = c.q * __builtin_amdgcn_rcpf((float)n);
    p = p + mk((float)fx, (float)fy, (float)fz) * s;
  }
  if (dr != 0.0f) p = grasp_pull<true>(p, tip, dr, vs.ks);
  if (!boundary) return with_w(p, p4.w);
  if (p.y < vs.ground) p.y = vs.ground;
  const f3 v = (p - xyz(x4)) * vs.inv_dt;
  *xs = with_w(p, x4.w);
  return with_w(predict(p, v, x4.w, vs.g, vs.dt, vs.damp1), p4.w);
}

template <int BLOCK, bool CARRY>
__device__ void share_fold_env(const ffsrl_batch& b, const ffsrl_params& prm, const ShareWs& ws, int el, int env,
                               int refit, unsigned long long* lds) {
  __shared__ int s_chunk, s_ok, s_last;
  const ffsrl_env_desc e = b.envs[env];
  const SharePlan pl = share_plan(e, prm, false);
  const int unit = pl.cc + pl.cv, I = prm.iterations, U = prm.substeps * I;
  uint32_t* ctr = ws.ctr_of(el, env);
  const int nv = e.nv, tid = threadIdx.x;
  const int64_t vb = e.vbase;
  const FrameCtx c = uniform(FrameCtx{lds, ldexpf(1.0f, -e.dp_exp), ldexpf(1.0f, e.dp_exp)});
  float4* Pl = reinterpret_cast<float4*>(lds + 2 * nv);  // positions of the unit this workgroup projects
  VertexStep vs;
  vs.g = mk(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
  vs.dt = prm.dt;
  vs.damp1 = 1.0f - prm.damping;
  vs.inv_dt = 1.0f / prm.dt;
  vs.ground = prm.ground;
  vs.ks = prm.drag_ks;
  const NHConst nhk = uniform(nh_const(prm.nh_vol_c, prm.nh_dev_c, prm.dt));
  const f3 tip = mk(b.tips[(env * 4 + 2) * 3 + 0], b.tips[(env * 4 + 2) * 3 + 1], b.tips[(env * 4 + 2) * 3 + 2]);
  const int64_t T = b.n_tet_total;
  const bool reset_lambda = !prm.lambda_legacy || prm.iterations > 1;
  const int nt = (prm.vol_mode != FFSRL_OFF || prm.nh_on) ? e.nt : 0;
  const int ne = prm.dist_mode != FFSRL_OFF ? e.ne : 0;
  const int nmax = nt > ne ? nt : ne;
  const float* X = b.x + 4 * vb;  // the env's rows of the model state
  const float* V = b.v + 4 * vb;
  auto sums = [&](int u) { return ws.fold_plane(u % 3, e); };
  auto pub = [&](int u) { return reinterpret_cast<float*>(ws.fold_plane(3 + u % 2, e)); };
  auto start = [&](int s) { return s == 0 ? X : reinterpret_cast<const float*>(ws.fold_plane(5 + s % 2, e)); };

  // positions of unit u of vertex i from the published state; *xs and
  // *boundary report a new substep's start position
  // positions of unit u of vertices lo .. hi-1 from the published state,
  // handed to out(i, position, new substep start).  Each thread takes K
  // vertices at a time and issues all their loads before any use (indices
  // clamped into range, only the hand-over guarded), so the loads of a group
  // are in flight together.
  constexpr int K = 4;
  auto for_unit = [&](int u, int lo, int hi, auto&& out) {
    const int last = hi - 1;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (u == 0) {
      for (int i0 = lo + tid; i0 < hi; i0 += K * BLOCK) {
        float4 x[K], v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int i = i0 + k * BLOCK < hi ? i0 + k * BLOCK : last;
          x[k] = ld4(X, i);
          v[k] = ld4(V, i);
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (i0 + k * BLOCK < hi)
            out(i0 + k * BLOCK, with_w(predict(xyz(x[k]), xyz(v[k]), x[k].w, vs.g, vs.dt, vs.damp1), x[k].w), zero);
      }
      return;
    }
    const int prev = u - 1;
    const unsigned long long* g = sums(prev);
    const float* pp = pub(prev);
    const float* st = start(prev / I);
    auto run = [&](auto boundary) {
      constexpr bool BND = decltype(boundary)::value;
      for (int i0 = lo + tid; i0 < hi; i0 += K * BLOCK) {
        float4 p4[K], x4[K];
        unsigned long long w0[K], w1[K];
        float dr[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int i = i0 + k * BLOCK < hi ? i0 + k * BLOCK : last;
          p4[k] = ld4c(pp, i);
          w0[k] = ldu64c(g, 2 * i);
          w1[k] = ldu64c(g, 2 * i + 1);
          dr[k] = b.drag[vb + i];
          x4[k] = BND ? ld4c(st, i) : zero;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (i0 + k * BLOCK >= hi) continue;
          float4 xs = zero;
          const float4 p = fold_advance(c, vs, tip, p4[k], w0[k], w1[k], dr[k], BND, x4[k], &xs);
          out(i0 + k * BLOCK, p, xs);
        }
      }
    };
    if (prev % I == I - 1)
      run(std::integral_constant<bool, true>());
    else
      run(std::integral_constant<bool, false>());
  };

  for (int i = tid; i < 2 * nv; i += BLOCK) lds[i] = 0ull;
  if (tid == 0) sat_flag() = 0;
  int pend_base = -1, pend = 0, pend_u = 0, lds_u = -1;
  for (;;) {
    __syncthreads();  // LDS and the shared words of the previous chunk are done with
    if (tid == 0) s_chunk = (int)atomicAdd(&ctr[0], 1u);
    __syncthreads();
    const int ci = s_chunk;
    int pass = 0, base = 0, local = 0;
    bool proj = false, fin = false;
    if (ci < pl.total) {
      pass = ci / unit;  // U for the final pass (its cv chunks follow the U units)
      base = pass * unit;
      local = ci - base;
      fin = pass == U;
      proj = !fin && local < pl.cc;
      if (!proj && !fin) local -= pl.cc;
    }
    if (pend && (ci >= pl.total || base != pend_base)) {
      // leave the pass: add the LDS sums to the unit's sums, then count the chunks
      unsigned long long* g = sums(pend_u);
      for (int i = tid; i < nv; i += BLOCK) {
        const unsigned long long w0 = lds[2 * i], w1 = lds[2 * i + 1];
        if (w0) {
          atomicAdd(&g[2 * i], w0);
          lds[2 * i] = 0ull;
        }
        if (w1) {
          atomicAdd(&g[2 * i + 1], w1);
          lds[2 * i + 1] = 0ull;
        }
      }
      wait_vmem();
      __syncthreads();
      if (tid == 0) atomicAdd(&ctr[1], (uint32_t)pend);
      pend = 0;
    }
    if (ci >= pl.total) break;
    if (tid == 0) s_ok = base == 0 || share_wait(ctr, (uint32_t)base);
    __syncthreads();
    if (!s_ok) {
      if (tid == 0) atomicOr(&ctr[2], 1u);
      break;
    }
    if (proj) {
      if (lds_u != pass) {  // the first projection chunk of this pass here: the unit's positions
        for_unit(pass, 0, nv, [&](int i, float4 p, float4) { Pl[i] = p; });
        __syncthreads();
        lds_u = pass;
      }
      const int jb = local * kShareItems, je = jb + kShareItems < nmax ? jb + kShareItems : nmax;
      project<BLOCK, CARRY, true>(b, prm, e, c, nhk, PosLds{Pl}, Items{b.tets16, b.edges16, 0, 0}, T, nt, ne, jb, je,
                                  reset_lambda && pass % I == 0);
      pend_base = base;
      pend_u = pass;
      pend += 1;
      continue;
    }
    const int v0 = local * kShareVerts, v1 = v0 + kShareVerts < nv ? v0 + kShareVerts : nv;
    if (!fin) {
      // publish unit `pass` (and a new substep's start), clear the sums of unit pass - 2
      float* P = pub(pass);
      float* Xs = pass > 0 && (pass - 1) % I == I - 1 ? const_cast<float*>(start(pass / I)) : nullptr;
      unsigned long long* old = pass >= 2 ? sums(pass - 2) : nullptr;
      for_unit(pass, v0, v1, [&](int i, float4 p, float4 xs) {
        st4(P, i, p);
        if (Xs) st4(Xs, i, xs);
        if (old) {
          old[2 * i] = 0ull;
          old[2 * i + 1] = 0ull;
        }
      });
    } else {
      // the last apply (+ grasp), ground clamp and integrate: the model's x, v, p
      const int last = U - 1;
      const unsigned long long* g = sums(last);
      const float* P = pub(last);
      const float* Xs = start(last / I);
      unsigned long long* g1 = sums(last);
      unsigned long long* g2 = U >= 2 ? sums(last - 1) : nullptr;
      for (int i = v0 + tid; i < v1; i += BLOCK) {
        int32_t fx, fy, fz, n;
        unpack2(ldu64c(g, 2 * i), &fx, &fy);
        unpack2(ldu64c(g, 2 * i + 1), &fz, &n);
        n = sum_count(n);
        const float4 p4 = ld4c(P, i);
        f3 p = xyz(p4);
        if (n > 0) {
          const float sc = c.q * __builtin_amdgcn_rcpf((float)n);
          p = p + mk((float)fx, (float)fy, (float)fz) * sc;
        }
        const float dr = b.drag[vb + i];
        if (dr != 0.0f) p = grasp_pull<true>(p, tip, dr, vs.ks);
        if (p.y < vs.ground) p.y = vs.ground;
        const float4 x4 = ld4c(Xs, i);
        const f3 v = (p - xyz(x4)) * vs.inv_dt;
        st4(b.x, vb + i, with_w(p, x4.w));
        st4(b.v, vb + i, with_w(v, 0.0f));
        st4(b.p, vb + i, with_w(p, p4.w));
        g1[2 * i] = 0ull;
        g1[2 * i + 1] = 0ull;
        if (g2) {
          g2[2 * i] = 0ull;
          g2[2 * i + 1] = 0ull;
        }
      }
    }
    wait_vmem();
    __syncthreads();
    if (tid == 0) s_last = atomicAdd(&ctr[1], 1u) + 1u == (uint32_t)pl.total;
    __syncthreads();
    if (refit && s_last) refit_env<BLOCK, true>(b, e);  // every chunk of the frame is done
  }
  __syncthreads();
  if (tid == 0 && sat_flag()) atomicOr(&b.status[env], (int32_t)FFSRL_STATUS_SATURATED);
}

// Hardware XCD id (0..15) -> dense index 0..nxcd-1 (-1: no XCD of that id).
struct XcdMap {
  int8_t index[16];
};

// FOLD (without COLLIDE only): envs of at most kFoldMaxVertices vertices take
// the folded form (ws.fold is set), the others the unfolded one.  The
// unfolded kernels carry no folded code, so their registers are unchanged.
template <int BLOCK, bool CARRY, bool COLLIDE, bool FOLD>
__global__ __launch_bounds__(BLOCK) void k_frame_shared(ffsrl_batch b, ffsrl_params prm, const int32_t* list, int n,
                                                        int refit, int nxcd, XcdMap map, ShareWs ws) {
  // float4-typed and named for this kernel, so the dynamic LDS starts 16-byte
  // aligned after the static LDS whatever that size is (block-scope extern
  // __shared__ arrays of one name are one entity, whose alignment is its
  // type's; misaligned 16-byte LDS accesses run far slower)
  extern __shared__ float4 lds_shared[];
  unsigned long long* lds = reinterpret_cast<unsigned long long*>(lds_shared);
  __shared__ int s_lastwg, s_incomplete;
  const int x = map.index[xcc_id() & 15];
  const int mine = (x >= 0 && x < nxcd && x < n) ? (n - 1 - x) / nxcd + 1 : 0;  // slots x, x + nxcd, ...
  const int rank = (int)blockIdx.x / nxcd;                           // spreads the XCD's workgroups
  for (int k = 0; k < mine; ++k) {
    const int el = x + nxcd * ((rank + k) % mine);
    const int env = list ? list[el] : el;
    if (FOLD && !COLLIDE && b.envs[env].nv <= kFoldMaxVertices)
      share_fold_env<BLOCK, CARRY>(b, prm, ws, el, env, refit, lds);
    else
      share_env<BLOCK, CARRY, COLLIDE>(b, prm, ws, el, env, refit, lds);
  }
  // The last workgroup out checks that every env's frame completed, clears
  // the global sums of any frame that did not (a workgroup that gave up
  // leaves partial sums), reports it, and resets the counters for the next
  // launch.  Every other workgroup has exited, so nothing else touches them.
  __syncthreads();
  uint32_t* exits = ws.ctr_of(0, list ? list[0] : 0) + 3;
  if (threadIdx.x == 0) s_lastwg = atomicAdd(exits, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_lastwg) return;
  // one thread per env: report an incomplete frame and reset its counters,
  // marking the frame (2 in its give-up word) for the clearing of its sums
  // by the whole workgroup below, which only runs when a frame is incomplete
  if (threadIdx.x == 0) s_incomplete = 0;
  __syncthreads();
  for (int el = threadIdx.x; el < n; el += BLOCK) {
    const int env = list ? list[el] : el;
    uint32_t* ctr = ws.ctr_of(el, env);
    const bool inc = atomicAdd(&ctr[1], 0u) != (uint32_t)share_plan(b.envs[env], prm, COLLIDE).total;
    if (inc) {
      atomicOr(&b.status[env], (int32_t)FFSRL_STATUS_INCOMPLETE);
      s_incomplete = 1;
    }
    atomicExch(&ctr[0], 0u);
    atomicExch(&ctr[1], 0u);
    atomicExch(&ctr[2], inc ? 2u : 0u);
  }
  __syncthreads();
  if (s_incomplete) {
    for (int el = 0; el < n; ++el) {
      const int env = list ? list[el] : el;
      const ffsrl_env_desc e = b.envs[env];
      uint32_t* ctr = ws.ctr_of(el, env);
      __syncthreads();
      if (threadIdx.x == 0) s_lastwg = atomicAdd(&ctr[2], 0u) == 2u;  // reused: this env's sums to clear
      __syncthreads();
      if (!s_lastwg) continue;
      if (FOLD && !COLLIDE && e.nv <= kFoldMaxVertices) {
        for (int k = 0; k < 3; ++k) {  // the three sum planes of the folded form
          unsigned long long* g = ws.fold_plane(k, e);
          for (int i = threadIdx.x; i < 2 * e.nv; i += BLOCK) atomicExch(&g[i], 0ull);
        }
      } else {
        unsigned long long* gacc = ws.acc_of(el, e);
        for (int i = threadIdx.x; i < 2 * e.nv; i += BLOCK) atomicExch(&gacc[i], 0ull);
      }
      if (threadIdx.x == 0) atomicExch(&ctr[2], 0u);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) atomicExch(exits, 0u);
}

// ---------------------------------------------------------------------------
// Partitioned frame (ffsrl_frame_parts): each env's constraints are cut into
// parts, contiguous ranges of its tets and of its edges in their current
// (Morton) order, so a part covers a compact region of the mesh; one
// workgroup runs one part for the whole frame.  A part's vertices (those its
// constraints touch, plus untouched vertices assigned to it) have part-local
// ids; their positions, substep-start positions, grasp weights and Jacobi sums
// stay in the workgroup's LDS for the frame, and the projection gathers from
// LDS.  A vertex held by several parts ("shared") is advanced by each of them
// from the same inputs with the same expressions, so the copies agree; its
// sums are exchanged through global planes: after projecting pass u (u =
// substep * iterations + iteration) a part adds its LDS sums of shared
// vertices into plane u % 3, counts its arrival, waits for every part of the
// env to arrive, and reads the complete sums back; a vertex of one part uses
// its LDS sums directly.  Integer sums do not depend on the order of the
// additions, so the frame equals the one-workgroup frame bit for bit.
//
// Plane reuse: the owner of a shared vertex clears its entry of plane
// (u - 1) % 3 after pass u's arrivals (every part read that plane in pass
// u - 1, before arriving at pass u); the plane is next added to in pass u + 2,
// after pass u + 1's arrivals, which include the owner's (issued after its
// clears completed).  One more arrival after the last pass lets the owners
// clear the last plane, so every frame leaves the planes at zero.
//
// Scheduling: list slot el belongs to XCD el % nxcd, so all parts of an env
// share one L2 (the coherence point, as in the shared frame), and the
// workgroups of an XCD claim its parts in slot order from a counter.  A part
// is claimed only by a running workgroup, after every part before it in the
// order; so every env whose parts are all claimed completes, and its
// workgroups then claim the next ones.  An env needs all its parts resident
// at once: at most as many parts as the XCD has CUs (checked at launch), and
// the launch should be the only persistent frame on the device (a partly
// resident launch can leave an env's parts unclaimed until its waits give up,
// which the frame reports as incomplete).  Waits are bounded as in the shared
// frame; the last workgroup out checks, cleans up and resets every counter.
struct PartsWs {
  const int32_t* parts;      // [n_parts][8]
  const int32_t* env_parts;  // int2 [n_env]
  const int32_t* verts;      // per part, its vertices (mesh-local id | flags)
  const uint64_t* t16;       // part-local packed tet ids
  const uint32_t* e16;       // part-local packed edge ids
  const uint64_t* f16;       // by part, its surface triangles' local ids (collision)
  unsigned long long* sums;  // [3][stride][2]
  int64_t stride;
  uint32_t* ctr;             // [n_env][4]: parts finished, pass arrivals, give-up flag, 0; then XCD claim
                             // counters [16], exits, and the claim table's first entry and size per XCD [16][2]
  int32_t* claims;           // per XCD, its parts in claim order: list slot << 6 | part
  int n_env;
  __device__ unsigned long long* plane(int k, int64_t vb) const { return sums + 2 * ((int64_t)k * stride + vb); }
  __device__ uint32_t* env_ctr(int env) const { return ctr + 4 * (int64_t)env; }
  __device__ uint32_t* xcd_ctr(int x) const { return ctr + 4 * (int64_t)n_env + x; }
  __device__ uint32_t* exits() const { return ctr + 4 * (int64_t)n_env + 16; }
  __device__ uint32_t* claim_info(int x) const { return ctr + 4 * (int64_t)n_env + 32 + 2 * x; }
};

// The claim table of a launch (one workgroup, before the frame on the same
// stream): XCD x's entries are the parts of its slots x, x + nxcd, ... in
// order, from claim_info(x)[0], claim_info(x)[1] of them, so a workgroup's
// k-th claim is one load.  Wavefront x counts XCD x's parts (a sweep of 64
// slots at a time, lanes summed by shuffles), the XCDs' first entries follow
// from the counts, and wavefront x writes its entries in a second sweep with
// an inclusive scan over its lanes.
__global__ __launch_bounds__(1024) void k_parts_claims(const int32_t* list, int n, int nxcd, PartsWs ws) {
  __shared__ int count[16], first[16];
  const int lane = threadIdx.x & 63, x = threadIdx.x >> 6;
  const int cnt = x < nxcd && x < n ? (n - 1 - x) / nxcd + 1 : 0;  // slots of XCD x
  auto parts_of = [&](int j) { return j < cnt ? ws.env_parts[2 * (list ? list[x + nxcd * j] : x + nxcd * j) + 1] : 0; };
  int total = 0;
  for (int j0 = 0; j0 < cnt; j0 += 64) {
    int v = parts_of(j0 + lane);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    total += v;
  }
  if (lane == 0 && x < 16) count[x] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    int base = 0;
    for (int k = 0; k < nxcd && k < 16; ++k) {
      first[k] = base;
      base += count[k];
    }
  }
  __syncthreads();
  if (x >= nxcd) return;
  int base = first[x];
  for (int j0 = 0; j0 < cnt; j0 += 64) {
    const int j = j0 + lane, np = parts_of(j);
    int incl = np;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off);
      if (lane >= off) incl += t;
    }
    const int el = x + nxcd * j;
    for (int q = 0; q < np; ++q) ws.claims[base + incl - np + q] = (el << 6) | q;
    base += __shfl(incl, 63);
  }
  if (lane == 0) {
    ws.claim_info(x)[0] = (uint32_t)first[x];
    ws.claim_info(x)[1] = (uint32_t)count[x];
  }
}

constexpr uint32_t kPartOwner = FFSRL_PART_OWNER, kPartShared = FFSRL_PART_SHARED;
constexpr uint32_t kPartIdMask = kPartShared - 1u;
// LDS per local vertex: sums 16 B, position 16 B, substep start 16 B, flags 4 B, grasp weight 4 B.
// 1 KB is left for the partitioned kernel's static LDS (544 B today: the
// paths' pass words, the saturation flag and the refit's group offsets;
// checked at launch).
constexpr int kPartStaticLds = 1024;
constexpr int kPartLocalBytes = 56;
constexpr int kPartMaxLocal = (kLdsBytes - kPartStaticLds) / kPartLocalBytes;
// An env in one part: sums and positions only (32 B per vertex).
constexpr int kSoloBytes = 32;
constexpr int kSoloMaxVertices = (kLdsBytes - kPartStaticLds) / kSoloBytes;

// An env held by one part: its vertex list is the identity and no vertex is
// shared, so this is the one-workgroup frame (frame_env) with the positions
// in LDS: the projection and the collision pass gather from LDS, the applies
// read and write LDS, and the substep-start positions, velocities and grasp
// weights stay in the model's arrays as there.  Each pass has no exchange and
// no wait; the model's p is written once, after the last pass.
template <bool CARRY, bool COLLIDE>
__device__ void solo_env(const ffsrl_batch& b, const ffsrl_params& prm, const PartsWs& ws, int env, int refit,
                         unsigned long long* lds) {
  constexpr int BLOCK = 1024;
  const ffsrl_env_desc e = b.envs[env];
  const int nv = e.nv, tid = threadIdx.x;
  const FrameCtx c = uniform(FrameCtx{lds, ldexpf(1.0f, -e.dp_exp), ldexpf(1.0f, e.dp_exp)});
  float4* Pl = reinterpret_cast<float4*>(lds + 2 * nv);
  const int64_t vb = e.vbase;
  float* X = b.x;
  float* Vv = b.v;
  VertexStep vs;
  vs.g = mk(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
  vs.dt = prm.dt;
  vs.damp1 = 1.0f - prm.damping;
  vs.inv_dt = 1.0f / prm.dt;
  vs.ground = prm.ground;
  vs.ks = prm.drag_ks;
  const NHConst nhk = uniform(nh_const(prm.nh_vol_c, prm.nh_dev_c, prm.dt));
  const f3 tip = mk(b.tips[(env * 4 + 2) * 3 + 0], b.tips[(env * 4 + 2) * 3 + 1], b.tips[(env * 4 + 2) * 3 + 2]);
  const int64_t T = b.n_tet_total;
  const bool reset_lambda = !prm.lambda_legacy || prm.iterations > 1;
  const int nt = (prm.vol_mode != FFSRL_OFF || prm.nh_on) ? e.nt : 0;
  const int ne = prm.dist_mode != FFSRL_OFF ? e.ne : 0;
  const int nmax = nt > ne ? nt : ne;
  const int I = prm.iterations;
  __shared__ int s_contacts;  // surface triangles in contact during the frame
  // ground clamp, integrate and the next substep's predict of vertex i
  auto finish = [&](int i, f3 p, float w, bool next_sub) {
    if (p.y < vs.ground) p.y = vs.ground;
    const float4 x4 = ld4(X, vb + i);
    const f3 v = (p - xyz(x4)) * vs.inv_dt;
    st4(X, vb + i, with_w(p, x4.w));
    if (next_sub)
      p = predict(p, v, x4.w, vs.g, vs.dt, vs.damp1);
    else
      st4(Vv, vb + i, with_w(v, 0.0f));
    Pl[i] = with_w(p, w);
  };

  // predict of substep 0; each thread notes which of its vertices (i = tid +
  // k * BLOCK, k < 5 since nv <= 5,088) are grasped
  uint32_t grasped = 0;
  for (int i = tid, k = 0; i < nv; i += BLOCK, ++k) {
    const float4 x = ld4(X, vb + i), v = ld4(Vv, vb + i);
    Pl[i] = with_w(predict(xyz(x), xyz(v), x.w, vs.g, vs.dt, vs.damp1), x.w);
    if (b.drag[vb + i] != 0.0f) grasped |= 1u << k;
    lds[2 * i] = 0ull;
    lds[2 * i + 1] = 0ull;
  }
  if (tid == 0) {
    sat_flag() = 0;
    s_contacts = 0;
  }
  __syncthreads();
  for (int s = 0; s < prm.substeps; ++s) {
    for (int it = 0; it < I; ++it) {
      project<BLOCK, CARRY, false>(b, prm, e, c, nhk, PosLds{Pl}, Items{b.tets16, b.edges16, 0, 0}, T, nt, ne, 0,
                                   nmax, reset_lambda && it == 0);
      __syncthreads();
      const bool last_it = it == I - 1, next_sub = s + 1 < prm.substeps;
      for (int i = tid, k = 0; i < nv; i += BLOCK, ++k) {
        const unsigned long long w0 = lds[2 * i], w1 = lds[2 * i + 1];
        lds[2 * i] = 0ull;
        lds[2 * i + 1] = 0ull;
        int32_t fx, fy, fz, n;
        unpack2(w0, &fx, &fy);
        unpack2(w1, &fz, &n);
        n = sum_count(n);
        const float4 p4 = Pl[i];
        f3 p = xyz(p4);
        if (n > 0) {
          const float sc = c.q * __builtin_amdgcn_rcpf((float)n);
          p = p + mk((float)fx, (float)fy, (float)fz) * sc;
        }
        if ((grasped >> k) & 1u) p = grasp_pull<true>(p, tip, b.drag[vb + i], vs.ks);
        if (last_it && !COLLIDE) {  // with collision, the integrate follows it
          if (p.y < vs.ground) p.y = vs.ground;
          const float4 x4 = ld4(X, vb + i);
          const f3 v = (p - xyz(x4)) * vs.inv_dt;
          st4(X, vb + i, with_w(p, x4.w));
          if (next_sub)
            p = predict(p, v, x4.w, vs.g, vs.dt, vs.damp1);
          else
            st4(Vv, vb + i, with_w(v, 0.0f));
        }
        Pl[i] = with_w(p, p4.w);
      }
      __syncthreads();
      if (COLLIDE) {
        // ---- instrument collision (collideLaparoscopeDO), summed in LDS ----
        collide_tris(b, prm, env, c, PosLds{Pl}, TrisGlobal{b.tris, (int64_t)e.fbase}, &s_contacts, BLOCK, 0, e.nf);
        __syncthreads();
        for (int i = tid; i < nv; i += BLOCK) {
          const unsigned long long w0 = lds[2 * i], w1 = lds[2 * i + 1];
          lds[2 * i] = 0ull;
          lds[2 * i + 1] = 0ull;
          if (!last_it && (w0 | w1) == 0ull) continue;
          int32_t fx, fy, fz, n;
          unpack2(w0, &fx, &fy);
          unpack2(w1, &fz, &n);
          (void)sum_count(n);  // collision terms carry no count, only saturation marks
          const float4 p4 = Pl[i];
          const f3 p = xyz(p4) + mk((float)fx, (float)fy, (float)fz) * c.q;
          if (last_it)
            finish(i, p, p4.w, next_sub);
          else
            Pl[i] = with_w(p, p4.w);
        }
        __syncthreads();
      }
    }
  }
  for (int i = tid; i < nv; i += BLOCK) st4(b.p, vb + i, Pl[i]);
  if (tid == 0) {
    if (COLLIDE && s_contacts) atomicAdd(&b.in_collision[env], s_contacts);  // counts accumulate until cleared
    atomicAdd(&ws.env_ctr(env)[0], 1u);  // the env's one part is done
    if (sat_flag()) atomicOr(&b.status[env], (int32_t)FFSRL_STATUS_SATURATED);
  }
  __syncthreads();
  if (refit) refit_env<BLOCK>(b, e);
}

// A part of a multi-part env.  Each pass (projection; with COLLIDE also the
// collision pass) sums into LDS, and the shared vertices' sums then cross one
// exchange: exchange x adds them into plane x % 3, arrives and waits for the
// env's np parts (np (x + 1) arrivals); every part applies the same totals to
// its copy of a shared vertex, and the owner clears plane (x - 1) % 3, which
// every part has read before arriving at exchange x.
template <bool CARRY, bool COLLIDE>
__device__ void part_env(const ffsrl_batch& b, const ffsrl_params& prm, const PartsWs& ws, int env, int pi, int np,
                         int refit, unsigned long long* lds) {
  constexpr int BLOCK = 1024;
  constexpr int E = COLLIDE ? 2 : 1;  // exchanges per pass
  __shared__ int s_ok, s_last, s_contacts;
  const ffsrl_env_desc e = b.envs[env];
  const int* pr = ws.parts + 8 * (int64_t)pi;
  const int t0 = pr[0], t1 = pr[1], e0 = pr[2], e1 = pr[3], voff = pr[4], nloc = pr[5];
  const int tid = threadIdx.x, I = prm.iterations, U = prm.substeps * I;
  const int64_t vb = e.vbase;
  const FrameCtx c = uniform(FrameCtx{lds, ldexpf(1.0f, -e.dp_exp), ldexpf(1.0f, e.dp_exp)});
  float4* Pl = reinterpret_cast<float4*>(lds + 2 * nloc);
  float4* Xl = Pl + nloc;
  uint32_t* Ml = reinterpret_cast<uint32_t*>(Xl + nloc);
  float* Dl = reinterpret_cast<float*>(Ml + nloc);
  VertexStep vs;
  vs.g = mk(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
  vs.dt = prm.dt;
  vs.damp1 = 1.0f - prm.damping;
  vs.inv_dt = 1.0f / prm.dt;
  vs.ground = prm.ground;
  vs.ks = prm.drag_ks;
  const NHConst nhk = uniform(nh_const(prm.nh_vol_c, prm.nh_dev_c, prm.dt));
  const f3 tip = mk(b.tips[(env * 4 + 2) * 3 + 0], b.tips[(env * 4 + 2) * 3 + 1], b.tips[(env * 4 + 2) * 3 + 2]);
  const int64_t T = b.n_tet_total;
  const bool reset_lambda = !prm.lambda_legacy || prm.iterations > 1;
  const int nt = (prm.vol_mode != FFSRL_OFF || prm.nh_on) ? t1 - t0 : 0;
  const int ne = prm.dist_mode != FFSRL_OFF ? e1 - e0 : 0;
  const int nmax = nt > ne ? nt : ne;
  const int X = E * U;  // exchanges in the frame
  uint32_t* ctr = ws.env_ctr(env);
  // this part's sums of its shared vertices into plane x % 3, then arrive and
  // wait for the env's parts; false when the wait gave up
  auto exchange = [&](int x) -> bool {
    unsigned long long* gs = ws.plane(x % 3, vb);
    for (int i = tid; i < nloc; i += BLOCK) {
      const uint32_t m = Ml[i];
      if (m & kPartShared) {
        const int g = (int)(m & kPartIdMask);
        const unsigned long long w0 = lds[2 * i], w1 = lds[2 * i + 1];
        if (w0) atomicAdd(&gs[2 * g], w0);
        if (w1) atomicAdd(&gs[2 * g + 1], w1);
      }
    }
    wait_vmem();
    __syncthreads();
    if (tid == 0) {
      atomicAdd(&ctr[1], 1u);
      s_ok = share_wait(ctr, (uint32_t)(np * (x + 1)));
    }
    __syncthreads();
    return s_ok != 0;
  };
  // the sums of local vertex i after an exchange into plane gs (LDS cleared
  // for the next pass); the owner of a shared vertex clears its entries of
  // the exchange's previous plane old (none before the first exchange)
  auto take = [&](int i, uint32_t m, const unsigned long long* gs, unsigned long long* old, unsigned long long* w0,
                  unsigned long long* w1) {
    *w0 = lds[2 * i];
    *w1 = lds[2 * i + 1];
    lds[2 * i] = 0ull;
    lds[2 * i + 1] = 0ull;
    if (m & kPartShared) {
      const int g = (int)(m & kPartIdMask);
      *w0 = ldu64c(gs, 2 * g);
      *w1 = ldu64c(gs, 2 * g + 1);
      if (old && (m & kPartOwner)) {
        old[2 * g] = 0ull;
        old[2 * g + 1] = 0ull;
      }
    }
  };
  // ground clamp, integrate and the next substep's predict of local vertex i
  auto finish = [&](int i, uint32_t m, f3 p, float w, bool next_sub) {
    if (p.y < vs.ground) p.y = vs.ground;
    const float4 x4 = Xl[i];
    const f3 v = (p - xyz(x4)) * vs.inv_dt;
    Xl[i] = with_w(p, x4.w);
    if (next_sub)
      p = predict(p, v, x4.w, vs.g, vs.dt, vs.damp1);
    else if (m & kPartOwner)
      st4(b.v, vb + (m & kPartIdMask), with_w(v, 0.0f));
    Pl[i] = with_w(p, w);
  };

  // the part's vertices: flags, state and the predict of substep 0
  for (int i = tid; i < nloc; i += BLOCK) {
    const uint32_t m = (uint32_t)ws.verts[voff + i];
    const int64_t g = vb + (m & kPartIdMask);
    const float4 x = ld4(b.x, g), v = ld4(b.v, g);
    Ml[i] = m;
    Xl[i] = x;
    Dl[i] = b.drag[g];
    Pl[i] = with_w(predict(xyz(x), xyz(v), x.w, vs.g, vs.dt, vs.damp1), x.w);
    lds[2 * i] = 0ull;
    lds[2 * i + 1] = 0ull;
  }
  if (tid == 0) {
    sat_flag() = 0;
    s_contacts = 0;
  }
  __syncthreads();
  bool ok = true;
  for (int u = 0; u < U; ++u) {
    const int s = u / I, itn = u - s * I;
    project<BLOCK, CARRY, false>(b, prm, e, c, nhk, PosLds{Pl}, Items{ws.t16, ws.e16, t0, e0}, T, nt, ne, 0, nmax,
                                 reset_lambda && itn == 0);
    __syncthreads();
    if (!exchange(E * u)) {
      ok = false;
      break;
    }
    // apply (+ grasp; at the end of a substep the ground clamp, integrate and next predict), as
    // the one-workgroup frame
    const bool last_it = itn == I - 1, next_sub = s + 1 < prm.substeps;
    const int x = E * u;
    const unsigned long long* gs = ws.plane(x % 3, vb);
    unsigned long long* old = x >= 1 ? ws.plane((x + 2) % 3, vb) : nullptr;
    for (int i = tid; i < nloc; i += BLOCK) {
      const uint32_t m = Ml[i];
      unsigned long long w0, w1;
      take(i, m, gs, old, &w0, &w1);
      int32_t fx, fy, fz, n;
      unpack2(w0, &fx, &fy);
      unpack2(w1, &fz, &n);
      n = sum_count(n);
      const float4 p4 = Pl[i];
      f3 p = xyz(p4);
      if (n > 0) {
        const float sc = c.q * __builtin_amdgcn_rcpf((float)n);
        p = p + mk((float)fx, (float)fy, (float)fz) * sc;
      }
      const float dr = Dl[i];
      if (dr != 0.0f) p = grasp_pull<true>(p, tip, dr, vs.ks);
      if (last_it && !COLLIDE)  // with collision, the integrate follows it
        finish(i, m, p, p4.w, next_sub);
      else
        Pl[i] = with_w(p, p4.w);
    }
    __syncthreads();
    if (COLLIDE) {
      // ---- instrument collision over the part's triangles, summed in LDS, then exchanged ----
      collide_tris(b, prm, env, c, PosLds{Pl}, TrisPacked{ws.f16, (int64_t)pr[6]}, &s_contacts, BLOCK, 0, pr[7]);
      __syncthreads();
      if (!exchange(x + 1)) {
        ok = false;
        break;
      }
      const unsigned long long* gc = ws.plane((x + 1) % 3, vb);
      unsigned long long* oc = ws.plane(x % 3, vb);
      for (int i = tid; i < nloc; i += BLOCK) {
        const uint32_t m = Ml[i];
        unsigned long long w0, w1;
        take(i, m, gc, oc, &w0, &w1);
        if (!last_it && (w0 | w1) == 0ull) continue;
        int32_t fx, fy, fz, n;
        unpack2(w0, &fx, &fy);
        unpack2(w1, &fz, &n);
        (void)sum_count(n);  // collision terms carry no count, only saturation marks
        const float4 p4 = Pl[i];
        const f3 p = xyz(p4) + mk((float)fx, (float)fy, (float)fz) * c.q;
        if (last_it)
          finish(i, m, p, p4.w, next_sub);
        else
          Pl[i] = with_w(p, p4.w);
      }
      __syncthreads();
    }
  }
  if (ok) {
    // the model's state of the vertices this part owns, then one more arrival
    for (int i = tid; i < nloc; i += BLOCK) {
      const uint32_t m = Ml[i];
      if (m & kPartOwner) {
        const int64_t g = vb + (m & kPartIdMask);
        st4(b.x, g, Xl[i]);
        st4(b.p, g, Pl[i]);
      }
    }
    wait_vmem();
    __syncthreads();
    if (tid == 0) {
      atomicAdd(&ctr[1], 1u);
      s_ok = share_wait(ctr, (uint32_t)(np * (X + 1)));
    }
    __syncthreads();
    ok = s_ok;
  }
  if (ok) {
    // every part has read the last plane: the owners clear it
    unsigned long long* last = ws.plane((X - 1) % 3, vb);
    for (int i = tid; i < nloc; i += BLOCK) {
      const uint32_t m = Ml[i];
      if ((m & kPartShared) && (m & kPartOwner)) {
        const int g = (int)(m & kPartIdMask);
        last[2 * g] = 0ull;
        last[2 * g + 1] = 0ull;
      }
    }
    wait_vmem();
    __syncthreads();
    if (tid == 0) {
      if (COLLIDE && s_contacts) atomicAdd(&b.in_collision[env], s_contacts);  // counts accumulate until cleared
      s_last = atomicAdd(&ctr[0], 1u) + 1u == (uint32_t)np;
    }
    __syncthreads();
    if (refit && s_last) refit_env<BLOCK, true>(b, e);  // every part's positions are stored
  } else if (tid == 0) {
    atomicOr(&ctr[2], 1u);  // the env's other waits give up too
  }
  __syncthreads();
  if (tid == 0 && sat_flag()) atomicOr(&b.status[env], (int32_t)FFSRL_STATUS_SATURATED);
}

template <bool CARRY, bool COLLIDE>
__global__ __launch_bounds__(1024) void k_frame_parts(ffsrl_batch b, ffsrl_params prm, const int32_t* list, int n,
                                                      int refit, int nxcd, XcdMap map, PartsWs ws) {
  // float4-typed and named for this kernel: the dynamic LDS starts 16-byte
  // aligned (the part's positions are float4s)
  extern __shared__ float4 lds_parts[];
  unsigned long long* lds = reinterpret_cast<unsigned long long*>(lds_parts);
  __shared__ int s_claim, s_lastwg, s_incomplete;
  const int x = map.index[xcc_id() & 15];
  if (x >= 0 && x < nxcd) {
    const uint32_t c0 = ws.claim_info(x)[0], cn = ws.claim_info(x)[1];
    for (;;) {
      if (threadIdx.x == 0) s_claim = (int)atomicAdd(ws.xcd_ctr(x), 1u);
      __syncthreads();
      const uint32_t k = (uint32_t)__builtin_amdgcn_readfirstlane(s_claim);
      __syncthreads();  // s_claim is written again by the next claim
      if (k >= cn) break;
      // the k-th part of this XCD's slots x, x + nxcd, ... in order
      const int entry = ws.claims[c0 + k], el = entry >> 6, part = entry & 63;
      const int env = list ? list[el] : el;
      const int np = ws.env_parts[2 * env + 1];
      if (np == 1)
        solo_env<CARRY, COLLIDE>(b, prm, ws, env, refit, lds);
      else
        part_env<CARRY, COLLIDE>(b, prm, ws, env, ws.env_parts[2 * env] + part, np, refit, lds);
      __syncthreads();  // LDS is reused by the next part
    }
  }
  // The last workgroup out checks that every env's frame completed, clears
  // the planes of any that did not (a part that gave up leaves partial
  // sums), reports it, and resets the counters for the next launch.
  __syncthreads();
  if (threadIdx.x == 0) s_lastwg = atomicAdd(ws.exits(), 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_lastwg) return;
  // one thread per env, then the planes of incomplete frames (marked 2 in
  // the give-up word) by the whole workgroup, as in the shared frame
  if (threadIdx.x == 0) s_incomplete = 0;
  __syncthreads();
  for (int el = threadIdx.x; el < n; el += 1024) {
    const int env = list ? list[el] : el;
    uint32_t* ctr = ws.env_ctr(env);
    const bool inc =
        atomicAdd(&ctr[0], 0u) != (uint32_t)ws.env_parts[2 * env + 1] || atomicAdd(&ctr[2], 0u) != 0u;
    if (inc) {
      atomicOr(&b.status[env], (int32_t)FFSRL_STATUS_INCOMPLETE);
      s_incomplete = 1;
    }
    atomicExch(&ctr[0], 0u);
    atomicExch(&ctr[1], 0u);
    atomicExch(&ctr[2], inc ? 2u : 0u);
  }
  __syncthreads();
  if (s_incomplete) {
    for (int el = 0; el < n; ++el) {
      const int env = list ? list[el] : el;
      const ffsrl_env_desc e = b.envs[env];
      uint32_t* ctr = ws.env_ctr(env);
      __syncthreads();
      if (threadIdx.x == 0) s_lastwg = atomicAdd(&ctr[2], 0u) == 2u;  // reused: this env's planes to clear
      __syncthreads();
      if (!s_lastwg) continue;
      for (int k = 0; k < 3; ++k) {
        unsigned long long* g = ws.plane(k, e.vbase);
        for (int i = threadIdx.x; i < 2 * e.nv; i += 1024) atomicExch(&g[i], 0ull);
      }
      if (threadIdx.x == 0) atomicExch(&ctr[2], 0u);
    }
  }
  __syncthreads();
  if (threadIdx.x < 16) atomicExch(ws.xcd_ctr(threadIdx.x), 0u);
  if (threadIdx.x == 0) atomicExch(ws.exits(), 0u);
}

__global__ void k_xcc_probe(int32_t* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

template <int BLOCK>
static int launch_shared(const ffsrl_batch& b, const ffsrl_params& prm, const int32_t* list, int n, int refit,
                         bool carry, int lds, int nxcd, const XcdMap& map, const ShareWs& ws, hipStream_t st) {
  const bool fold = ws.fold != nullptr;
  auto kern = prm.collide ? (carry ? k_frame_shared<BLOCK, true, true, false> : k_frame_shared<BLOCK, false, true, false>)
              : fold ? (carry ? k_frame_shared<BLOCK, true, false, true> : k_frame_shared<BLOCK, false, false, true>)
                     : (carry ? k_frame_shared<BLOCK, true, false, false> : k_frame_shared<BLOCK, false, false, false>);
  FFSRL_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  int dev = 0, cus = 0, per_cu = 0;
  FFSRL_HIP(hipGetDevice(&dev));
  FFSRL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  FFSRL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), BLOCK, lds));
  const int grid = std::max(nxcd, cus * std::max(1, per_cu));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), lds, st, b, prm, list, n, refit, nxcd, map, ws);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

static int check_batch(const ffsrl_batch* b, const ffsrl_params* prm) {
  if (!b || !b->envs || !b->x || !b->p || !b->v || !b->drag || !b->tips)
    return fail(FFSRL_E_ARG, "batch has a NULL state or env array");
  if (b->n_env < 0) return fail(FFSRL_E_ARG, "n_env < 0");
  if (prm) {
    if (prm->substeps < 1 || prm->iterations < 1) return fail(FFSRL_E_ARG, "substeps and iterations must be >= 1");
    if (!(prm->dt > 0.0f)) return fail(FFSRL_E_ARG, "dt must be > 0");
    if (prm->dist_mode < 0 || prm->dist_mode > 2 || prm->vol_mode < 0 || prm->vol_mode > 2)
      return fail(FFSRL_E_ARG, "constraint mode out of range");
    if ((prm->vol_mode || prm->nh_on) && (!b->tets || !b->tet_rest_volume))
      return fail(FFSRL_E_ARG, "tet constraints enabled without tet arrays");
    if (prm->nh_on && !b->tet_inv_rest) return fail(FFSRL_E_ARG, "Neo-Hookean enabled without inverse rest matrices");
    if (prm->dist_mode && (!b->edges || !b->edge_rest)) return fail(FFSRL_E_ARG, "distance enabled without edge arrays");
  }
  return 0;
}

static bool carries_lambda(const ffsrl_params& prm) {
  const bool xpbd = prm.dist_mode == FFSRL_XPBD || prm.vol_mode == FFSRL_XPBD;
  return xpbd && (prm.iterations > 1 || prm.lambda_legacy);
}

// ---------------------------------------------------------------- plain form

__global__ void k_predict(ffsrl_batch b, ffsrl_params prm, int64_t vb, int nv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  float4 x = ld4(b.x, vb + i), v = ld4(b.v, vb + i);
  f3 g = mk(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
  st4(b.p, vb + i, with_w(predict(xyz(x), xyz(v), x.w, g, prm.dt, 1.0f - prm.damping), x.w));
}

// Items of one plain launch: those of env `env` (off NULL: item = thread),
// or of envs eb .. eb+n-1 in one launch (the reference's launch model: one
// launch per kernel over every environment), where thread t is item
// t - off[k] of env eb + k, off[] holding the prefix sums of the envs' item
// counts (off[k] <= t < off[k+1], found by binary search).
struct ItemMap {
  const int64_t* off;
  int eb, n, env;
};
__device__ __forceinline__ bool map_item(const ItemMap& m, int64_t t, int* env, int* local) {
  if (!m.off) {
    *env = m.env;
    *local = (int)t;
    return true;
  }
  if (t >= m.off[m.n]) return false;
  int lo = 0, hi = m.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.off[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  *env = m.eb + lo;
  *local = (int)(t - m.off[lo]);
  return true;
}
__device__ __forceinline__ int64_t gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

__device__ __forceinline__ void g_add(const ffsrl_batch& b, int64_t V, int64_t i, f3 d) {
  unsafeAtomicAdd(&b.dp[i], d.x);
  unsafeAtomicAdd(&b.dp[V + i], d.y);
  unsafeAtomicAdd(&b.dp[2 * V + i], d.z);
}

__global__ void k_volume(ffsrl_batch b, ffsrl_params prm, ffsrl_env_desc e1, ItemMap m, int64_t V) {
  int env, t;
  if (!map_item(m, gtid(), &env, &t)) return;
  const ffsrl_env_desc e = m.off ? b.envs[env] : e1;
  if (t >= e.nt) return;
  const int64_t ti = e.tbase + t, vb = e.vbase;
  const int4 q = ldi4(b.tets, ti);
  float4 Pq[4] = {ld4(b.p, vb + q.x), ld4(b.p, vb + q.y), ld4(b.p, vb + q.z), ld4(b.p, vb + q.w)};
  f3 d[4];
  int ok;
  if (prm.vol_mode == FFSRL_PBD) {
    ok = volume_pbd<false>(Pq, b.tet_rest_volume[ti], prm.vol_k, d);
  } else {
    float lam = b.lam_t ? b.lam_t[e.lam_tbase + t] : 0.0f;
    ok = volume_xpbd<false>(Pq, b.tet_rest_volume[ti], prm.vol_k, &lam, d);
    if (b.lam_t && ok) b.lam_t[e.lam_tbase + t] = lam;
  }
  if (!ok) return;
  const int ids[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    g_add(b, V, vb + ids[k], d[k]);
    atomicAdd(&b.cnt[vb + ids[k]], 1);
  }
}

__global__ void k_neohookean(ffsrl_batch b, ffsrl_params prm, ffsrl_env_desc e1, ItemMap m, int64_t V) {
  int env, t;
  if (!map_item(m, gtid(), &env, &t)) return;
  const ffsrl_env_desc e = m.off ? b.envs[env] : e1;
  if (t >= e.nt) return;
  const int64_t ti = e.tbase + t, vb = e.vbase, T = b.n_tet_total;
  const int4 q = ldi4(b.tets, ti);
  float4 Pq[4] = {ld4(b.p, vb + q.x), ld4(b.p, vb + q.y), ld4(b.p, vb + q.z), ld4(b.p, vb + q.w)};
  float M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = b.tet_inv_rest[k * T + ti];
  f3 d[4];
  if (!neohookean<false>(Pq, b.tet_rest_volume[ti], M, nh_const(prm.nh_vol_c, prm.nh_dev_c, prm.dt), d)) return;
  const int ids[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    g_add(b, V, vb + ids[k], d[k]);
    atomicAdd(&b.cnt[vb + ids[k]], 1);
  }
}

__global__ void k_distance(ffsrl_batch b, ffsrl_params prm, ffsrl_env_desc e1, ItemMap m, int64_t V) {
  int env, k;
  if (!map_item(m, gtid(), &env, &k)) return;
  const ffsrl_env_desc e = m.off ? b.envs[env] : e1;
  if (k >= e.ne) return;
  const int64_t ei = e.ebase + k, vb = e.vbase;
  const int2 q = ldi2(b.edges, ei);
  const float4 A = ld4(b.p, vb + q.x), B = ld4(b.p, vb + q.y);
  f3 d;
  int ok;
  if (prm.dist_mode == FFSRL_PBD) {
    ok = distance_pbd<false>(A, B, b.edge_rest[ei], prm.dist_k, &d);
  } else {
    float lam = b.lam_e ? b.lam_e[e.lam_ebase + k] : 0.0f;
    ok = distance_xpbd<false>(A, B, b.edge_rest[ei], prm.dist_k, &lam, &d);
    if (b.lam_e && ok) b.lam_e[e.lam_ebase + k] = lam;
  }
  if (!ok) return;
  g_add(b, V, vb + q.x, -(d * A.w));
  g_add(b, V, vb + q.y, d * B.w);
  atomicAdd(&b.cnt[vb + q.x], 1);
  atomicAdd(&b.cnt[vb + q.y], 1);
}

__global__ void k_apply(ffsrl_batch b, int64_t vb, int nv, int64_t V) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int64_t g = vb + i;
  const int n = b.cnt[g];
  if (n > 0) {
    float4 p = ld4(b.p, g);
    const float N = (float)n;
    f3 q = xyz(p) + mk(b.dp[g] / N, b.dp[V + g] / N, b.dp[2 * V + g] / N);
    st4(b.p, g, with_w(q, p.w));
  }
  b.dp[g] = 0.0f;
  b.dp[V + g] = 0.0f;
  b.dp[2 * V + g] = 0.0f;
  b.cnt[g] = 0;
}

__global__ void k_grasp(ffsrl_batch b, ffsrl_params prm, int64_t vb1, int nv1, ItemMap m) {
  int env, i;
  if (!map_item(m, gtid(), &env, &i)) return;
  const int64_t vb = m.off ? (int64_t)b.envs[env].vbase : vb1;
  if (!m.off && i >= nv1) return;
  const float dr = b.drag[vb + i];
  if (dr == 0.0f) return;
  const float* t = b.tips + (env * 4 + 2) * 3;
  float4 p = ld4(b.p, vb + i);
  st4(b.p, vb + i, with_w(grasp_pull<false>(xyz(p), mk(t[0], t[1], t[2]), dr, prm.drag_ks), p.w));
}

// collideLaparoscopeDO (integrator.py:981-1050), plain form: one thread per
// surface triangle of env `env`; responses summed into dp (no count), then
// k_collide_apply subtracts them, so every triangle reads the positions of
// the start of the pass.
__global__ void k_collide(ffsrl_batch b, ffsrl_params prm, ffsrl_env_desc e1, ItemMap m, int64_t V) {
  int env, f;
  if (!map_item(m, gtid(), &env, &f)) return;
  const ffsrl_env_desc e = m.off ? b.envs[env] : e1;
  if (f >= e.nf) return;
  const int64_t vb = e.vbase;
  const int4 t = ldi4(b.tris, (int64_t)e.fbase + f);
  const float4 A = ld4(b.p, vb + t.x), B = ld4(b.p, vb + t.y), Cc = ld4(b.p, vb + t.z);
  const f3 pA = xyz(A), pB = xyz(B), pC = xyz(Cc);
  const f3 centre = ((pA + pB) + pC) * 0.3333f;
  const f3 n = normalize(cross(pB - pA, pC - pA));
  f3 r[3] = {mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f)};
  float len = 0.0f;
  for (int i = 0; i < 3; ++i) {
    const int ci = env * 4 + i;
    const f3 tip = mk(b.tips[3 * ci], b.tips[3 * ci + 1], b.tips[3 * ci + 2]);
    const f3 base = mk(b.caps_base[3 * ci], b.caps_base[3 * ci + 1], b.caps_base[3 * ci + 2]);
    if (!tri_meets_box(capsule_box(tip, base, b.caps_radius[ci]), pA, pB, pC)) continue;  // zero response
    len = len + capsule_response(tip, base, b.caps_radius[ci], n, centre, pA, pB, pC, prm.contact_k, r);
  }
  if (!(len > 0.0f)) return;
  atomicAdd(&b.in_collision[env], 1);
  if (A.w > 0.0f) g_add(b, V, vb + t.x, r[0]);
  if (B.w > 0.0f) g_add(b, V, vb + t.y, r[1]);
  if (Cc.w > 0.0f) g_add(b, V, vb + t.z, r[2]);
}

__global__ void k_collide_apply(ffsrl_batch b, int64_t vb, int nv, int64_t V) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int64_t g = vb + i;
  const float4 p = ld4(b.p, g);
  st4(b.p, g, with_w(xyz(p) - mk(b.dp[g], b.dp[V + g], b.dp[2 * V + g]), p.w));
  b.dp[g] = 0.0f;
  b.dp[V + g] = 0.0f;
  b.dp[2 * V + g] = 0.0f;
}

__global__ void k_bounds_integrate(ffsrl_batch b, ffsrl_params prm, int64_t vb, int nv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  float4 p = ld4(b.p, vb + i), x = ld4(b.x, vb + i);
  if (p.y < prm.ground) p.y = prm.ground;  // bounds (integrator.py:12-21)
  st4(b.p, vb + i, p);
  const f3 v = (xyz(p) - xyz(x)) * (1.0f / prm.dt);  // PBDStep (integrator.py:23-38)
  st4(b.x, vb + i, with_w(xyz(p), x.w));
  st4(b.v, vb + i, with_w(v, 0.0f));
}

__global__ void k_zero(float* a, int64_t base, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[base + i] = 0.0f;
}

// One thread per env: the components of delta that would take the drag
// point (point 3) out of [lo, hi] are dropped (transfromCartesianActions-
// KernelInWorkspace, modelDO.py:97-154), and the four points move by the
// rest.
__global__ void k_move_tools(ffsrl_batch b, const float* delta, const float* lo, const float* hi) {
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= b.n_env) return;
  float d[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float moved = b.tips[env * 12 + 9 + k] + delta[env * 3 + k];
    d[k] = ((lo && moved < lo[k]) || (hi && moved > hi[k])) ? 0.0f : delta[env * 3 + k];
  }
  for (int p = 0; p < 4; ++p) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = env * 12 + p * 3 + k;
      const float t0 = b.tips[i], t = t0 + d[k];
      b.tips[i] = t;
      // capsule bases follow their tips (the instrument is translated rigidly)
      if (b.caps_base) b.caps_base[i] += t - t0;
    }
  }
}

__global__ void k_grasp_select(ffsrl_batch b, int env, float radius, const int32_t* state,
                               const int32_t* update) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const ffsrl_env_desc e = b.envs[env];
  if (i >= e.nv) return;
  if (update[env * 3] == 0) return;
  const int64_t g = (int64_t)e.vbase + i;
  if (state[env * 3] == 1) {
    const float* t = b.tips + (env * 4 + 3) * 3;
    const float len = length(xyz(ld4(b.x, g)) - mk(t[0], t[1], t[2]));
    b.drag[g] = len < radius ? 1.0f : 0.0f;
  } else {
    b.drag[g] = 0.0f;
  }
}

__global__ void k_remap(float* dst, const float* src, const int32_t* map, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  st4(dst, map[2 * i], ld4(src, map[2 * i + 1]));
}

}  // namespace ffsrl

using namespace ffsrl;

namespace {
// The plain form launches per environment from the host, so it reads the
// caller's host mirror of the env table (ffsrl_batch.host_envs).
int plain_envs(const ffsrl_batch* b, int begin, int end, const ffsrl_env_desc** out, int64_t* vtotal) {
  if (begin < 0 || end > b->n_env || begin > end)
    return fail(FFSRL_E_ARG, "env range [%d,%d) outside [0,%d)", begin, end, b->n_env);
  if (!b->host_envs) return fail(FFSRL_E_ARG, "the plain form needs host_envs (host mirror of the env table)");
  int64_t v = 0;
  for (int i = 0; i < b->n_env; ++i) v = std::max<int64_t>(v, (int64_t)b->host_envs[i].vbase + b->host_envs[i].nv);
  *out = b->host_envs;
  *vtotal = v;
  return 0;
}

inline unsigned nblk(int n, int t) { return (unsigned)((n + t - 1) / t); }
}  // namespace

extern "C" {

int ffsrl_abi_version(void) { return FFSRL_ABI_VERSION; }
const char* ffsrl_last_error(void) { return g_err.c_str(); }
int ffsrl_frame_max_vertices(void) { return kMaxFrameVertices; }

// Argument checks shared by the batched frame entry points.
static int frame_checks(const ffsrl_batch* b, const ffsrl_params* prm, int32_t n, int32_t refit) {
  if (!prm) return fail(FFSRL_E_ARG, "params NULL");
  int rc = check_batch(b, prm);
  if (rc) return rc;
  if (b->max_nv < 0 || b->max_nv > kMaxFrameVertices)
    return fail(FFSRL_E_LIMIT, "max_nv %d exceeds the LDS capacity of the batched frame (%d vertices)", b->max_nv,
                kMaxFrameVertices);
  const bool carry = carries_lambda(*prm);
  if (carry && ((prm->dist_mode == FFSRL_XPBD && !b->lam_e) || (prm->vol_mode == FFSRL_XPBD && !b->lam_t)))
    return fail(FFSRL_E_ARG, "XPBD with iterations > 1 needs lambda arrays");
  if (refit && (!b->tris || !b->level_nodes || !b->level_info || !b->level_off || !b->bvh_lo || !b->bvh_hi))
    return fail(FFSRL_E_ARG, "refit requested without BVH arrays");
  if ((prm->vol_mode || prm->nh_on) && !b->tets16) return fail(FFSRL_E_ARG, "batched frame needs packed tet ids");
  if (prm->dist_mode && !b->edges16) return fail(FFSRL_E_ARG, "batched frame needs packed edge ids");
  if (prm->collide && (!b->tris || !b->caps_base || !b->caps_radius || !b->in_collision))
    return fail(FFSRL_E_ARG, "collision needs triangles, capsule bases / radii and the in_collision counters");
  if (n < 0 || n > b->n_env) return fail(FFSRL_E_ARG, "%d environments listed, batch has %d", n, b->n_env);
  if (!b->status) return fail(FFSRL_E_ARG, "the batched frames need the per-env status array");
  if (b->host_envs) {
    // a vertex's terms of one pass must fit the 32-bit fixed-point sums
    for (int i = 0; i < b->n_env; ++i)
      if (b->host_envs[i].max_terms > FFSRL_MAX_TERMS)
        return fail(FFSRL_E_LIMIT, "env %d: a vertex receives %d Jacobi terms in one pass (batched frame limit %d);"
                    " use the plain form", i, b->host_envs[i].max_terms, FFSRL_MAX_TERMS);
  }
  return 0;
}

static int frame(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* envs, int32_t n, int32_t block,
                 int32_t refit, uint32_t* sched, void* stream) {
  const int rc = frame_checks(b, prm, n, refit);
  if (rc) return rc;
  if (n == 0) return 0;
  const bool carry = carries_lambda(*prm);
  const int lds = lds_bytes_for(b->max_nv < 1 ? 1 : b->max_nv);
  hipStream_t st = as_stream(stream);
  switch (block) {
    case 256: return launch_frame<256>(*b, *prm, envs, n, refit, carry, lds, sched, st);
    case 512: return launch_frame<512>(*b, *prm, envs, n, refit, carry, lds, sched, st);
    case 1024: return launch_frame<1024>(*b, *prm, envs, n, refit, carry, lds, sched, st);
    default: return fail(FFSRL_E_ARG, "block must be 256, 512 or 1024 (got %d)", block);
  }
}

int ffsrl_frame(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* order, int32_t block,
                int32_t refit, void* stream) {
  if (!b) return fail(FFSRL_E_ARG, "batch NULL");
  return frame(b, prm, order, b->n_env, block, refit, b->sched, stream);
}

int ffsrl_frame_envs(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* envs, int32_t n, int32_t block,
                     int32_t refit, void* stream) {
  if (n > 0 && !envs) return fail(FFSRL_E_ARG, "env list NULL");
  return frame(b, prm, envs, n, block, refit, nullptr, stream);
}

// XCD count of each device, probed once with one workgroup per CU; the
// hardware ids seen are numbered densely (XcdMap) for the shared frame.
static std::mutex g_xcd_mu;
static int g_xcd[64];
static XcdMap g_xcd_map[64];

int ffsrl_xcd_count(void) {
  int dev = 0, cus = 0;
  FFSRL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FFSRL_E_LIMIT, "device %d", dev);
  std::lock_guard<std::mutex> lock(g_xcd_mu);
  if (g_xcd[dev] > 0) return g_xcd[dev];
  FFSRL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int32_t* d = nullptr;
  FFSRL_HIP(hipMalloc(&d, sizeof(int32_t) * cus));
  std::vector<int32_t> h(cus, -1);
  hipLaunchKernelGGL(k_xcc_probe, dim3(cus), dim3(64), 0, 0, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(h.data(), d, sizeof(int32_t) * cus, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  FFSRL_HIP(e);
  bool seen[16] = {};
  for (int v : h) {
    if (v < 0 || v >= 16) return fail(FFSRL_E_LIMIT, "XCD id %d outside 0..15", v);
    seen[v] = true;
  }
  // dense numbering of the ids seen (a partitioned device may report one id, e.g. 3)
  XcdMap map;
  int n = 0;
  for (int i = 0; i < 16; ++i) map.index[i] = seen[i] ? (int8_t)n++ : (int8_t)-1;
  g_xcd_map[dev] = map;
  g_xcd[dev] = n;
  return n;
}

// ---------------------------------------------------------------------------
// No-allocation shared frame (ffsrl_frame_shared_noalloc): the global sums
// and chunk counters live in `scratch`, one copy per device for the life of
// the process, indexed by list slot: slot el owns sums for max_nv vertices
// at el * max_nv and counter row el.  Every frame leaves the sums and the
// counters it used at zero (an apply chunk clears the sums it reads; the
// last workgroup out resets the counters and clears the sums of a frame
// that did not complete), and the array is cleared once per device at its
// first use, so every call starts from zeros.
constexpr int64_t kScratchAccWords = (32ll << 20) / 8;  // 32 MiB of global sums (2 words per vertex)
constexpr int kScratchSlots = 4096;                    // list slots (counter rows of 4 words)
constexpr int64_t kScratchWords = kScratchAccWords + kScratchSlots * 4 / 2;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

__global__ void k_fill_u64(unsigned long long* a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

// Device address of `scratch` on the current device (fetched once per
// device, with that device current; the array's name in host code is a
// host-side stand-in, never a device address).
static std::mutex g_scratch_mu;
static void* g_scratch[64];

static int scratch_of(int dev, hipStream_t st, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    FFSRL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(FFSRL_E_STATE, "no device address for the frame scratch");
    // cleared once, by a fill kernel on the caller's stream (ordered before the frame)
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((kScratchWords + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long*>(ptr), kScratchWords);
    FFSRL_HIP(hipGetLastError());
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return 0;
}

int ffsrl_frame_shared_noalloc(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* envs, int32_t n,
                               int32_t block, int32_t refit, void* stream) {
  int rc = frame_checks(b, prm, n, refit);
  if (rc) return rc;
  if (n == 0) return 0;
  const int64_t slot_nv = b->max_nv < 1 ? 1 : b->max_nv;
  if (n > kScratchSlots || (int64_t)n * slot_nv * 2 > kScratchAccWords)
    return fail(FFSRL_E_LIMIT, "no-allocation shared frame: %d envs x %lld vertices exceed its workspace (%d envs, %lld "
                "vertex slots)", n, (long long)slot_nv, kScratchSlots, (long long)(kScratchAccWords / 2));
  int dev = 0;
  FFSRL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FFSRL_E_LIMIT, "device %d", dev);
  int nxcd;
  XcdMap map;
  {
    std::lock_guard<std::mutex> lock(g_xcd_mu);
    nxcd = g_xcd[dev];
    map = g_xcd_map[dev];
  }
  // the XCD probe allocates, so it is not run here
  if (nxcd <= 0) return fail(FFSRL_E_STATE, "call ffsrl_xcd_count() once on this device before the no-allocation frame");
  hipStream_t st = as_stream(stream);
  void* ws_ptr = nullptr;
  if ((rc = scratch_of(dev, st, &ws_ptr))) return rc;
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(ws_ptr);
  const ShareWs ws{acc, reinterpret_cast<uint32_t*>(acc + kScratchAccWords), 1, (int)slot_nv};
  const bool carry = carries_lambda(*prm);
  const int lds = lds_bytes_for((int)slot_nv);
  switch (block) {
    case 256: return launch_shared<256>(*b, *prm, envs, n, refit, carry, lds, nxcd, map, ws, st);
    case 512: return launch_shared<512>(*b, *prm, envs, n, refit, carry, lds, nxcd, map, ws, st);
    case 1024: return launch_shared<1024>(*b, *prm, envs, n, refit, carry, lds, nxcd, map, ws, st);
    default: return fail(FFSRL_E_ARG, "block must be 256, 512 or 1024 (got %d)", block);
  }
}

int ffsrl_frame_shared(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* envs, int32_t n, int32_t block,
                       int32_t refit, void* stream) {
  int rc = frame_checks(b, prm, n, refit);
  if (rc) return rc;
  if (!b->split_acc || !b->split_ctr) return fail(FFSRL_E_ARG, "shared frame needs split_acc and split_ctr");
  if (n == 0) return 0;
  const int nxcd = ffsrl_xcd_count();
  if (nxcd <= 0) return nxcd < 0 ? nxcd : fail(FFSRL_E_LIMIT, "no XCD found");
  int dev = 0;
  FFSRL_HIP(hipGetDevice(&dev));
  XcdMap map;
  {
    std::lock_guard<std::mutex> lock(g_xcd_mu);
    map = g_xcd_map[dev];
  }
  const bool carry = carries_lambda(*prm);
  const int max_nv = b->max_nv < 1 ? 1 : b->max_nv;
  int lds = lds_bytes_for(max_nv);
  const bool fold = b->fold_ws && !prm->collide;
  if (fold) {
    // every env with vbase + nv inside the planes (checked from the host mirror when given)
    if (b->fold_stride < 1) return fail(FFSRL_E_ARG, "fold_ws needs fold_stride");
    if (b->host_envs)
      for (int i = 0; i < b->n_env; ++i)
        if ((int64_t)b->host_envs[i].vbase + b->host_envs[i].nv > b->fold_stride)
          return fail(FFSRL_E_ARG, "fold_stride %lld < vbase + nv of env %d", (long long)b->fold_stride, i);
    lds = std::max(lds, 32 * std::min(max_nv, kFoldMaxVertices));  // positions + sums of a folded env
  }
  hipStream_t st = as_stream(stream);
  const ShareWs ws{reinterpret_cast<unsigned long long*>(b->split_acc), b->split_ctr, 0, 0,
                   fold ? reinterpret_cast<unsigned long long*>(b->fold_ws) : nullptr, b->fold_stride};
  switch (block) {
    case 256: return launch_shared<256>(*b, *prm, envs, n, refit, carry, lds, nxcd, map, ws, st);
    case 512: return launch_shared<512>(*b, *prm, envs, n, refit, carry, lds, nxcd, map, ws, st);
    case 1024: return launch_shared<1024>(*b, *prm, envs, n, refit, carry, lds, nxcd, map, ws, st);
    default: return fail(FFSRL_E_ARG, "block must be 256, 512 or 1024 (got %d)", block);
  }
}

int ffsrl_parts_max_local(void) { return kPartMaxLocal; }
int ffsrl_parts_max_solo(void) { return kSoloMaxVertices; }

int ffsrl_frame_parts(const ffsrl_batch* b, const ffsrl_params* prm, const ffsrl_parts* pt, const int32_t* envs,
                      int32_t n, int32_t refit, void* stream) {
  if (!prm) return fail(FFSRL_E_ARG, "params NULL");
  int rc = check_batch(b, prm);
  if (rc) return rc;
  if (!pt) return fail(FFSRL_E_ARG, "parts NULL");
  if (prm->collide && (!b->tris || !b->caps_base || !b->caps_radius || !b->in_collision || !pt->tris16))
    return fail(FFSRL_E_ARG, "collision needs triangles, capsule bases / radii, the in_collision counters and the "
                             "parts' triangles");
  const bool carry = carries_lambda(*prm);
  if (carry && ((prm->dist_mode == FFSRL_XPBD && !b->lam_e) || (prm->vol_mode == FFSRL_XPBD && !b->lam_t)))
    return fail(FFSRL_E_ARG, "XPBD with iterations > 1 needs lambda arrays");
  if (refit && (!b->tris || !b->level_nodes || !b->level_info || !b->level_off || !b->bvh_lo || !b->bvh_hi))
    return fail(FFSRL_E_ARG, "refit requested without BVH arrays");
  if ((prm->vol_mode || prm->nh_on) && (!pt->tets16 || !b->tets16))
    return fail(FFSRL_E_ARG, "partitioned frame needs part and packed tet ids");
  if (prm->dist_mode && (!pt->edges16 || !b->edges16))
    return fail(FFSRL_E_ARG, "partitioned frame needs part and packed edge ids");
  if (!pt->parts || !pt->env_parts || !pt->verts || !pt->sums || !pt->ctr || !pt->claims || !pt->host_env_parts)
    return fail(FFSRL_E_ARG, "parts: NULL table, sums, counters or claims");
  if (pt->max_local < 0 || pt->max_local > kPartMaxLocal)
    return fail(FFSRL_E_LIMIT, "a part of %d local vertices (LDS holds %d)", pt->max_local, kPartMaxLocal);
  if (pt->max_solo < 0 || pt->max_solo > kSoloMaxVertices)
    return fail(FFSRL_E_LIMIT, "an env of %d vertices in one part (LDS holds %d)", pt->max_solo, kSoloMaxVertices);
  if (pt->max_local < 1 && pt->max_solo < 1) return fail(FFSRL_E_ARG, "parts: no part sizes given");
  if (n < 0 || n > b->n_env) return fail(FFSRL_E_ARG, "%d environments listed, batch has %d", n, b->n_env);
  if (!b->status) return fail(FFSRL_E_ARG, "the batched frames need the per-env status array");
  if (n == 0) return 0;
  const int nxcd = ffsrl_xcd_count();
  if (nxcd <= 0) return nxcd < 0 ? nxcd : fail(FFSRL_E_LIMIT, "no XCD found");
  int dev = 0, cus = 0, per_cu = 0;
  FFSRL_HIP(hipGetDevice(&dev));
  FFSRL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  XcdMap map;
  {
    std::lock_guard<std::mutex> lock(g_xcd_mu);
    map = g_xcd_map[dev];
  }
  const int per_xcd = std::min(cus / nxcd, 63);
  int64_t total = 0;
  for (int i = 0; i < b->n_env; ++i) {
    const int np = pt->host_env_parts[2 * i + 1];
    total += np;
    if (np < 1 || np > per_xcd)
      return fail(FFSRL_E_LIMIT, "env %d has %d parts; they must run at once on one XCD (%d CUs)", i, np, per_xcd);
    if (b->host_envs) {
      const ffsrl_env_desc& h = b->host_envs[i];
      if (np == 1 && h.nv > pt->max_solo)
        return fail(FFSRL_E_ARG, "env %d of %d vertices in one part, max_solo %d", i, h.nv, pt->max_solo);
      if (h.max_terms > FFSRL_MAX_TERMS)
        return fail(FFSRL_E_LIMIT, "env %d: a vertex receives %d Jacobi terms in one pass (limit %d)", i,
                    h.max_terms, FFSRL_MAX_TERMS);
      if ((int64_t)h.vbase + h.nv > pt->stride)
        return fail(FFSRL_E_ARG, "parts stride %lld < vbase + nv of env %d", (long long)pt->stride, i);
    }
  }
  if (total > pt->n_claims) return fail(FFSRL_E_ARG, "parts: %lld claim entries for %lld parts", (long long)pt->n_claims,
                                        (long long)total);
  if (n >= (1 << 25)) return fail(FFSRL_E_LIMIT, "%d environments listed (claim entries hold 2^25)", n);
  const int lds = std::max(kPartLocalBytes * pt->max_local, kSoloBytes * pt->max_solo);
  auto kern = prm->collide ? (carry ? k_frame_parts<true, true> : k_frame_parts<false, true>)
                           : (carry ? k_frame_parts<true, false> : k_frame_parts<false, false>);
  hipFuncAttributes fa;
  FFSRL_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kern)));
  if ((int)fa.sharedSizeBytes > kPartStaticLds)
    return fail(FFSRL_E_LIMIT, "partitioned frame: %d B of static LDS exceed the %d B reserved", (int)fa.sharedSizeBytes,
                kPartStaticLds);
  FFSRL_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  FFSRL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), 1024, lds));
  const int grid = std::max(nxcd, cus * std::max(1, per_cu));
  const PartsWs ws{pt->parts,
                   pt->env_parts,
                   pt->verts,
                   pt->tets16,
                   pt->edges16,
                   pt->tris16,
                   reinterpret_cast<unsigned long long*>(pt->sums),
                   pt->stride,
                   pt->ctr,
                   pt->claims,
                   b->n_env};
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_parts_claims, dim3(1), dim3(1024), 0, st, envs, n, nxcd, ws);
  FFSRL_HIP(hipGetLastError());
  hipLaunchKernelGGL(kern, dim3(grid), dim3(1024), lds, st, *b, *prm, envs, n, refit, nxcd, map, ws);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_predict(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].nv) continue;
    hipLaunchKernelGGL(k_predict, dim3(nblk(h[i].nv, 256)), dim3(256), 0, st, *b, *prm, (int64_t)h[i].vbase, h[i].nv);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_project_volume(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  if (prm->vol_mode == FFSRL_OFF) return 0;
  if (!b->dp || !b->cnt) return fail(FFSRL_E_ARG, "plain form needs dp/cnt scratch");
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].nt) continue;
    hipLaunchKernelGGL(k_volume, dim3(nblk(h[i].nt, 256)), dim3(256), 0, st, *b, *prm, h[i], ItemMap{nullptr, 0, 0, i},
                       V);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_project_neohookean(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  if (!prm->nh_on) return 0;
  if (!b->dp || !b->cnt) return fail(FFSRL_E_ARG, "plain form needs dp/cnt scratch");
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].nt) continue;
    hipLaunchKernelGGL(k_neohookean, dim3(nblk(h[i].nt, 256)), dim3(256), 0, st, *b, *prm, h[i],
                       ItemMap{nullptr, 0, 0, i}, V);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_project_distance(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  if (prm->dist_mode == FFSRL_OFF) return 0;
  if (!b->dp || !b->cnt) return fail(FFSRL_E_ARG, "plain form needs dp/cnt scratch");
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].ne) continue;
    hipLaunchKernelGGL(k_distance, dim3(nblk(h[i].ne, 256)), dim3(256), 0, st, *b, *prm, h[i], ItemMap{nullptr, 0, 0, i},
                       V);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_apply(const ffsrl_batch* b, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, nullptr);
  if (rc) return rc;
  if (!b->dp || !b->cnt) return fail(FFSRL_E_ARG, "plain form needs dp/cnt scratch");
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].nv) continue;
    hipLaunchKernelGGL(k_apply, dim3(nblk(h[i].nv, 256)), dim3(256), 0, st, *b, (int64_t)h[i].vbase, h[i].nv, V);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_grasp(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].nv) continue;
    hipLaunchKernelGGL(k_grasp, dim3(nblk(h[i].nv, 256)), dim3(256), 0, st, *b, *prm, (int64_t)h[i].vbase, h[i].nv,
                       ItemMap{nullptr, 0, 0, i});
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_collide(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  if (!b->dp) return fail(FFSRL_E_ARG, "plain form needs dp/cnt scratch");
  if (!b->tris || !b->caps_base || !b->caps_radius || !b->in_collision)
    return fail(FFSRL_E_ARG, "collision needs triangles, capsule bases / radii and the in_collision counters");
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].nv || !h[i].nf) continue;
    hipLaunchKernelGGL(k_collide, dim3(nblk(h[i].nf, 256)), dim3(256), 0, st, *b, *prm, h[i], ItemMap{nullptr, 0, 0, i}, V);
    hipLaunchKernelGGL(k_collide_apply, dim3(nblk(h[i].nv, 256)), dim3(256), 0, st, *b, (int64_t)h[i].vbase, h[i].nv,
                       V);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_bounds_integrate(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (!h[i].nv) continue;
    hipLaunchKernelGGL(k_bounds_integrate, dim3(nblk(h[i].nv, 256)), dim3(256), 0, st, *b, *prm, (int64_t)h[i].vbase,
                       h[i].nv);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_zero_lambdas(const ffsrl_batch* b, int32_t eb, int32_t ee, void* stream) {
  int rc = check_batch(b, nullptr);
  if (rc) return rc;
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  for (int i = eb; i < ee; ++i) {
    if (b->lam_e && h[i].ne)
      hipLaunchKernelGGL(k_zero, dim3(nblk(h[i].ne, 256)), dim3(256), 0, st, b->lam_e, (int64_t)h[i].lam_ebase, h[i].ne);
    if (b->lam_t && h[i].nt)
      hipLaunchKernelGGL(k_zero, dim3(nblk(h[i].nt, 256)), dim3(256), 0, st, b->lam_t, (int64_t)h[i].lam_tbase, h[i].nt);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// The reference's launch model (ffsrl_frame_do): every plain kernel launched
// once per substep / iteration over all envs of the range, envs found in the
// prefix sums of their item counts (ItemMap); the refit one launch per refit
// group over all envs.  The prefix sums live in device memory kept per env
// range and uploaded again only when the host mirror's counts change.
}  // extern "C"

namespace {

__global__ void k_refit_group_do(ffsrl_batch b, ItemMap m, int L) {
  int env, k;
  if (!map_item(m, gtid(), &env, &k)) return;
  const ffsrl_env_desc e = b.envs[env];
  refit_node(b, e, b.level_off[e.lvlbase + L] + k);
}

struct DoMaps {
  std::vector<int64_t> host;  // [4 + groups][n + 1]: vertices, tets, edges, triangles, refit groups
  int64_t* dev = nullptr;
  size_t cap = 0;
};
std::mutex g_do_mu;
DoMaps g_do;

// Prefix sums of envs eb .. ee-1 (host mirror), uploaded when they changed.
int do_maps(const ffsrl_batch* b, int eb, int ee, int groups, const int64_t** dev) {
  const int n = ee - eb;
  std::vector<int64_t> h((size_t)(4 + groups) * (n + 1), 0);
  for (int k = 0; k < n; ++k) {
    const ffsrl_env_desc& e = b->host_envs[eb + k];
    const int64_t cnt[4] = {e.nv, e.nt, e.ne, e.nf};
    for (int a = 0; a < 4; ++a) h[(size_t)a * (n + 1) + k + 1] = h[(size_t)a * (n + 1) + k] + cnt[a];
    for (int L = 0; L < groups; ++L) {
      const int64_t c = L < e.nlvl ? b->host_level_off[e.lvlbase + L + 1] - b->host_level_off[e.lvlbase + L] : 0;
      const size_t r = (size_t)(4 + L) * (n + 1);
      h[r + k + 1] = h[r + k] + c;
    }
  }
  std::lock_guard<std::mutex> lock(g_do_mu);
  if (h != g_do.host) {
    // earlier frames (on any stream) may still read the buffer
    FFSRL_HIP(hipDeviceSynchronize());
    if (h.size() > g_do.cap) {
      if (g_do.dev) (void)hipFree(g_do.dev);
      g_do.dev = nullptr;
      g_do.cap = 0;
      FFSRL_HIP(hipMalloc(&g_do.dev, h.size() * sizeof(int64_t)));
      g_do.cap = h.size();
    }
    FFSRL_HIP(hipMemcpy(g_do.dev, h.data(), h.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    g_do.host.swap(h);
  }
  *dev = g_do.dev;
  return 0;
}

}  // namespace

extern "C" {

int ffsrl_frame_do(const ffsrl_batch* b, const ffsrl_params* prm, int32_t eb, int32_t ee, int32_t refit,
                   void* stream) {
  int rc = check_batch(b, prm);
  if (rc) return rc;
  if (!b->dp || !b->cnt) return fail(FFSRL_E_ARG, "the plain form needs dp/cnt scratch");
  const ffsrl_env_desc* h;
  int64_t V;
  if ((rc = plain_envs(b, eb, ee, &h, &V))) return rc;
  if (eb == ee) return 0;
  if (refit && (!b->level_off || !b->level_nodes || !b->level_info || !b->bvh_lo || !b->bvh_hi || !b->tris ||
                !b->host_level_off))
    return fail(FFSRL_E_ARG, "frame_do: the refit needs the BVH arrays and host_level_off");
  if (prm->collide && (!b->tris || !b->caps_base || !b->caps_radius || !b->in_collision))
    return fail(FFSRL_E_ARG, "frame_do: collision needs triangles, capsules and the in_collision counters");
  // the envs' vertices (and carried lambdas) must be contiguous: one range per launch
  for (int i = eb + 1; i < ee; ++i) {
    if ((int64_t)h[i].vbase != (int64_t)h[i - 1].vbase + h[i - 1].nv)
      return fail(FFSRL_E_ARG, "frame_do: vertices of envs %d and %d are not contiguous", i - 1, i);
    if (b->lam_t && (int64_t)h[i].lam_tbase != (int64_t)h[i - 1].lam_tbase + h[i - 1].nt)
      return fail(FFSRL_E_ARG, "frame_do: tet lambdas of envs %d and %d are not contiguous", i - 1, i);
    if (b->lam_e && (int64_t)h[i].lam_ebase != (int64_t)h[i - 1].lam_ebase + h[i - 1].ne)
      return fail(FFSRL_E_ARG, "frame_do: edge lambdas of envs %d and %d are not contiguous", i - 1, i);
  }
  const int n = ee - eb;
  int groups = 0;
  for (int i = eb; i < ee; ++i) groups = std::max(groups, (int)h[i].nlvl);
  if (!refit) groups = 0;
  const int64_t* off = nullptr;
  if ((rc = do_maps(b, eb, ee, groups, &off))) return rc;
  std::vector<int64_t> tot(4 + groups);
  for (int a = 0; a < 4 + groups; ++a) {
    int64_t t = 0;
    for (int i = eb; i < ee; ++i) {
      const ffsrl_env_desc& e = h[i];
      if (a < 4) {
        t += a == 0 ? e.nv : a == 1 ? e.nt : a == 2 ? e.ne : e.nf;
      } else {
        const int L = a - 4;
        if (L < e.nlvl) t += b->host_level_off[e.lvlbase + L + 1] - b->host_level_off[e.lvlbase + L];
      }
    }
    tot[a] = t;
  }
  auto map = [&](int a) { return ItemMap{off + (size_t)a * (n + 1), eb, n, 0}; };
  auto blocks = [](int64_t items) { return dim3((unsigned)((items + 255) / 256)); };
  const int64_t vb0 = h[eb].vbase;
  const int nvt = (int)tot[0];
  hipStream_t st = as_stream(stream);
  const bool carry = carries_lambda(*prm);
  const bool reset = carry && (!prm->lambda_legacy || prm->iterations > 1);
  for (int s = 0; s < prm->substeps; ++s) {
    if (nvt) hipLaunchKernelGGL(k_predict, blocks(nvt), dim3(256), 0, st, *b, *prm, vb0, nvt);
    if (reset) {
      if (b->lam_e && tot[2])
        hipLaunchKernelGGL(k_zero, blocks(tot[2]), dim3(256), 0, st, b->lam_e, (int64_t)h[eb].lam_ebase, (int)tot[2]);
      if (b->lam_t && tot[1])
        hipLaunchKernelGGL(k_zero, blocks(tot[1]), dim3(256), 0, st, b->lam_t, (int64_t)h[eb].lam_tbase, (int)tot[1]);
    }
    for (int it = 0; it < prm->iterations; ++it) {
      if (prm->vol_mode != FFSRL_OFF && tot[1])
        hipLaunchKernelGGL(k_volume, blocks(tot[1]), dim3(256), 0, st, *b, *prm, h[eb], map(1), V);
      if (prm->dist_mode != FFSRL_OFF && tot[2])
        hipLaunchKernelGGL(k_distance, blocks(tot[2]), dim3(256), 0, st, *b, *prm, h[eb], map(2), V);
      if (prm->nh_on && tot[1])
        hipLaunchKernelGGL(k_neohookean, blocks(tot[1]), dim3(256), 0, st, *b, *prm, h[eb], map(1), V);
      if (nvt) {
        hipLaunchKernelGGL(k_apply, blocks(nvt), dim3(256), 0, st, *b, vb0, nvt, V);
        hipLaunchKernelGGL(k_grasp, blocks(nvt), dim3(256), 0, st, *b, *prm, vb0, nvt, map(0));
      }
      if (prm->collide && tot[3] && nvt) {
        hipLaunchKernelGGL(k_collide, blocks(tot[3]), dim3(256), 0, st, *b, *prm, h[eb], map(3), V);
        hipLaunchKernelGGL(k_collide_apply, blocks(nvt), dim3(256), 0, st, *b, vb0, nvt, V);
      }
    }
    if (nvt) hipLaunchKernelGGL(k_bounds_integrate, blocks(nvt), dim3(256), 0, st, *b, *prm, vb0, nvt);
  }
  for (int L = 0; L < groups; ++L)
    if (tot[4 + L]) hipLaunchKernelGGL(k_refit_group_do, blocks(tot[4 + L]), dim3(256), 0, st, *b, map(4 + L), L);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_move_tools(const ffsrl_batch* b, const float* delta, const float* lo, const float* hi, void* stream) {
  if (!b || !b->tips || !delta) return fail(FFSRL_E_ARG, "tips and delta must be non-NULL");
  if (b->n_env == 0) return 0;
  hipLaunchKernelGGL(k_move_tools, dim3(nblk(b->n_env, 256)), dim3(256), 0, as_stream(stream), *b, delta, lo, hi);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_grasp_select(const ffsrl_batch* b, float radius, const int32_t* state, const int32_t* update,
                       void* stream) {
  int rc = check_batch(b, nullptr);
  if (rc) return rc;
  if (!state || !update) return fail(FFSRL_E_ARG, "state/update NULL");
  const ffsrl_env_desc* h;
  int64_t V;
  hipStream_t st = as_stream(stream);
  if ((rc = plain_envs(b, 0, b->n_env, &h, &V))) return rc;
  for (int i = 0; i < b->n_env; ++i) {
    if (!h[i].nv) continue;
    hipLaunchKernelGGL(k_grasp_select, dim3(nblk(h[i].nv, 256)), dim3(256), 0, st, *b, i, radius, state, update);
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_remap(float* dst, const float* src, const int32_t* map, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (!dst || !src || !map))) return fail(FFSRL_E_ARG, "remap arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_remap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), dst, src, map, n);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_graph_begin(void* stream) {
  FFSRL_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  return 0;
}

int ffsrl_graph_end(void* stream, void** graph_exec) {
  if (!graph_exec) return fail(FFSRL_E_ARG, "graph_exec NULL");
  hipGraph_t g = nullptr;
  FFSRL_HIP(hipStreamEndCapture(as_stream(stream), &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
  *graph_exec = ex;
  return 0;
}

int ffsrl_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return fail(FFSRL_E_ARG, "graph_exec NULL");
  FFSRL_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), as_stream(stream)));
  return 0;
}

int ffsrl_graph_abort(void* stream) {
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(as_stream(stream), &g);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
  return 0;
}

int ffsrl_graph_destroy(void* graph_exec) {
  if (!graph_exec) return 0;
  FFSRL_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec)));
  return 0;
}

}  // extern "C"
