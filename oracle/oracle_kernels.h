/*
 * TEST INFRASTRUCTURE ONLY - the oracle.  CPU restatement of the reference's
 * per-step kernels.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline.
 *
 * This header is included twice by oracle.c, once with REAL=float (SFX=_f32)
 * and once with REAL=double (SFX=_f64).  Every kernel runs its "threads" in
 * serial tid order, so the float atomics of the reference accumulate in tid
 * order (Warp's CPU execution model).  Arrays use the reference's layout:
 * AoS vec3 (x,y,z per vertex), interleaved int32 index arrays, global ids.
 *
 * Builtin arithmetic follows warp-lang 0.10.1 as restated here (Warp is absent
 * from this machine, so these choices are "parity unpinned" against Warp
 * itself and pinned instead by closed forms and brute force in tests/):
 *   length(v) = sqrt((x*x + y*y) + z*z)
 *   normalize(v) = v / length(v) component-wise if length > 0 else 0
 *   min(a,b) = a < b ? a : b,  max(a,b) = a > b ? a : b
 *   mat33(c0,c1,c2) takes COLUMNS; M[i] is row i; mul() sums k = 0,1,2
 *   determinant(M) = row0 . (row1 x row2)
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)
#define R(x) ((REAL)(x))

typedef struct { REAL x, y, z; } FN(vec3_t);
#define V3 FN(vec3_t)

static inline V3 FN(v3)(REAL x, REAL y, REAL z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 FN(ld)(const REAL* a, long i) { return FN(v3)(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
static inline void FN(st)(REAL* a, long i, V3 v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }
static inline V3 FN(add)(V3 a, V3 b) { return FN(v3)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 FN(sub)(V3 a, V3 b) { return FN(v3)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 FN(neg)(V3 a) { return FN(v3)(-a.x, -a.y, -a.z); }
static inline V3 FN(mul)(V3 a, REAL s) { return FN(v3)(a.x * s, a.y * s, a.z * s); }
static inline V3 FN(divs)(V3 a, REAL s) { return FN(v3)(a.x / s, a.y / s, a.z / s); }
static inline REAL FN(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 FN(cross)(V3 a, V3 b) {
  return FN(v3)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline REAL FN(sqrt_)(REAL x) { return (REAL)SQRT(x); }
static inline REAL FN(length)(V3 a) { return FN(sqrt_)(FN(dot)(a, a)); }
static inline V3 FN(normalize)(V3 a) {
  REAL l = FN(length)(a);
  if (l > R(0)) return FN(divs)(a, l);
  return FN(v3)(R(0), R(0), R(0));
}
static inline REAL FN(fmin_)(REAL a, REAL b) { return a < b ? a : b; }
static inline REAL FN(fmax_)(REAL a, REAL b) { return a > b ? a : b; }
static inline V3 FN(vmin)(V3 a, V3 b) { return FN(v3)(FN(fmin_)(a.x, b.x), FN(fmin_)(a.y, b.y), FN(fmin_)(a.z, b.z)); }
static inline V3 FN(vmax)(V3 a, V3 b) { return FN(v3)(FN(fmax_)(a.x, b.x), FN(fmax_)(a.y, b.y), FN(fmax_)(a.z, b.z)); }
static inline REAL FN(absr)(REAL a) { return a < R(0) ? -a : a; }

/* FLOAT_EPSILON = sys.float_info.epsilon (integrator.py:8). */
#define FEPS R(2.220446049250313e-16)

/* gravity (integrator.py:40-63) */
void FN(oracle_gravity)(long n, const REAL* vertex, REAL* pred, const REAL* vel, const REAL* invMass,
                        const REAL* g, REAL damp, REAL dt) {
  V3 G = FN(v3)(g[0], g[1], g[2]);
  REAL d = R(1.0) - damp;
  for (long tid = 0; tid < n; ++tid) {
    V3 x = FN(ld)(vertex, tid), v = FN(ld)(vel, tid);
    REAL w = invMass[tid];
    v = FN(add)(v, FN(mul)(FN(mul)(FN(mul)(G, w), dt), d));
    if (w < FEPS) v = FN(v3)(R(0), R(0), R(0));
    FN(st)(pred, tid, FN(add)(x, FN(mul)(v, dt)));
  }
}

/* distanceConstraints, PBD form (integrator.py:65-109) */
void FN(oracle_distance_pbd)(long nE, const REAL* pred, REAL* dP, int* cnt, const int* edge,
                             const REAL* rest, const REAL* invMass, const REAL* active, REAL kS) {
  for (long tid = 0; tid < nE; ++tid) {
    REAL act = active ? active[tid] : R(1);
    if (act == R(0)) continue;
    int a = edge[2 * tid], b = edge[2 * tid + 1];
    V3 pa = FN(ld)(pred, a), pb = FN(ld)(pred, b);
    REAL r = rest[tid];
    if (r < FEPS) continue;
    V3 dir = FN(sub)(pb, pa);
    REAL len = FN(length)(dir);
    if (len < FEPS) continue;
    REAL w = invMass[a] + invMass[b];
    if (w < FEPS) continue;
    REAL C = len - r;
    V3 e = FN(divs)(FN(mul)(FN(mul)(FN(mul)(FN(divs)(dir, len), -C), kS), act), w);
    FN(st)(dP, a, FN(sub)(FN(ld)(dP, a), FN(mul)(e, invMass[a])));
    FN(st)(dP, b, FN(add)(FN(ld)(dP, b), FN(mul)(e, invMass[b])));
    cnt[a] += 1; cnt[b] += 1;
  }
}

/* distanceConstraintsXPBD (integrator.py:191-240); alpha is compliance/dt^2 */
void FN(oracle_distance_xpbd)(long nE, const REAL* pred, REAL* dP, REAL* lambdas, int* cnt,
                              const int* edge, const REAL* rest, const REAL* invMass,
                              const REAL* active, REAL alpha) {
  for (long tid = 0; tid < nE; ++tid) {
    if (active && active[tid] == R(0)) continue;
    int a = edge[2 * tid], b = edge[2 * tid + 1];
    V3 pa = FN(ld)(pred, a), pb = FN(ld)(pred, b);
    REAL r = rest[tid];
    if (r < FEPS) continue;
    REAL lam = lambdas[tid];
    V3 dir = FN(sub)(pb, pa);
    REAL len = FN(length)(dir);
    if (len < FEPS) continue;
    REAL w = invMass[a] + invMass[b];
    if (w < FEPS) continue;
    REAL C = len - r;
    REAL dl = -(C + alpha * lam) / (w + alpha);
    lambdas[tid] = lam + dl;
    V3 e = FN(mul)(FN(normalize)(dir), dl);
    FN(st)(dP, a, FN(sub)(FN(ld)(dP, a), FN(mul)(e, invMass[a])));
    FN(st)(dP, b, FN(add)(FN(ld)(dP, b), FN(mul)(e, invMass[b])));
    cnt[a] += 1; cnt[b] += 1;
  }
}

/* shared part of the two volume kernels: q vectors and w */
static inline void FN(vol_terms)(const REAL* pred, const REAL* invMass, const int* t, V3 q[4], REAL m[4],
                                 REAL* six_vol, REAL* w) {
  V3 A = FN(ld)(pred, t[0]), B = FN(ld)(pred, t[1]), C = FN(ld)(pred, t[2]), D = FN(ld)(pred, t[3]);
  V3 p1 = FN(sub)(B, A), p2 = FN(sub)(C, A), p3 = FN(sub)(D, A);
  q[2] = FN(cross)(p3, p1);
  q[1] = FN(cross)(p2, p3);
  q[3] = FN(cross)(p1, p2);
  q[0] = FN(sub)(FN(sub)(FN(neg)(q[1]), q[2]), q[3]);
  for (int i = 0; i < 4; ++i) m[i] = invMass[t[i]];
  *six_vol = FN(dot)(FN(cross)(p1, p2), p3);
  *w = m[0] * FN(dot)(q[0], q[0]) + m[1] * FN(dot)(q[1], q[1]) + m[2] * FN(dot)(q[2], q[2]) +
       m[3] * FN(dot)(q[3], q[3]);
}

static inline void FN(scatter4)(REAL* dP, int* cnt, const int* t, const V3 d[4]) {
  for (int i = 0; i < 4; ++i) FN(st)(dP, t[i], FN(add)(FN(ld)(dP, t[i]), d[i]));
  for (int i = 0; i < 4; ++i) cnt[t[i]] += 1;
}

/* volumeConstraints, PBD form (integrator.py:444-504) */
void FN(oracle_volume_pbd)(long nT, const REAL* pred, REAL* dP, int* cnt, const int* tet,
                           const REAL* restVol, const REAL* invMass, const REAL* active, REAL kS) {
  for (long tid = 0; tid < nT; ++tid) {
    REAL act = active ? active[tid] : R(1);
    if (!(act == R(1))) continue;
    const int* t = tet + 4 * tid;
    V3 q[4], d[4]; REAL m[4], sv, w;
    FN(vol_terms)(pred, invMass, t, q, m, &sv, &w);
    REAL volume = FN(absr)(sv) / R(6.0);
    if (FN(absr)(w) < FEPS) continue;
    REAL C = restVol[tid] - volume;
    REAL dl = -kS * act * C / w;
    for (int i = 0; i < 4; ++i) d[i] = FN(mul)(FN(mul)(q[i], dl), m[i]);
    FN(scatter4)(dP, cnt, t, d);
  }
}

/* volumeConstraintsXPBD (integrator.py:507-576) */
void FN(oracle_volume_xpbd)(long nT, const REAL* pred, REAL* dP, REAL* lambdas, int* cnt,
                            const int* tet, const REAL* restVol, const REAL* invMass,
                            const REAL* active, REAL alpha) {
  for (long tid = 0; tid < nT; ++tid) {
    if (active && !(active[tid] == R(1))) continue;
    const int* t = tet + 4 * tid;
    V3 q[4], d[4]; REAL m[4], sv, w;
    REAL lam = lambdas[tid];
    FN(vol_terms)(pred, invMass, t, q, m, &sv, &w);
    REAL volume = sv / R(6.0);
    if (FN(absr)(w) < FEPS) continue;
    REAL C = volume - restVol[tid];
    REAL dl = -(C + alpha * lam) / (w + alpha);
    lambdas[tid] = lam + dl;
    for (int i = 0; i < 4; ++i) d[i] = FN(mul)(FN(mul)(q[i], dl), m[i]);
    FN(scatter4)(dP, cnt, t, d);
  }
}

/* 3x3 helpers, row-major m[3*i+j] */
static inline void FN(mat_cols)(REAL* m, V3 c0, V3 c1, V3 c2) {
  m[0] = c0.x; m[1] = c1.x; m[2] = c2.x;
  m[3] = c0.y; m[4] = c1.y; m[5] = c2.y;
  m[6] = c0.z; m[7] = c1.z; m[8] = c2.z;
}
static inline void FN(mat_mul)(REAL* o, const REAL* a, const REAL* b) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      REAL s = R(0);
      for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
      o[3 * i + j] = s;
    }
}
static inline void FN(mat_T)(REAL* o, const REAL* a) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[3 * j + i] = a[3 * i + j];
}
static inline V3 FN(row)(const REAL* m, int i) { return FN(v3)(m[3 * i], m[3 * i + 1], m[3 * i + 2]); }

/* Constraint values and gradient rows of neoHookeanConstraints
 * (integrator.py:601-683) for the tet positions A..D:
 *   hydrostatic C_H = det F - 1 - invD/invV with rows of transpose(mat33(F1xF2,
 *   F2xF0, F0xF1) * M^T); deviatoric C_D = |F|_F with rows of
 *   transpose(F * M^T * (1/|F|_F)); g0 = -g1 - g2 - g3 in both. */
typedef struct { REAL CH, CD; V3 gH[4], gD[4]; } FN(nh_terms_t);

static FN(nh_terms_t) FN(nh_terms)(V3 A, V3 B, V3 C3, V3 D, const REAL* M, REAL volC, REAL devC) {
  FN(nh_terms_t) o;
  V3 p1 = FN(sub)(B, A), p2 = FN(sub)(C3, A), p3 = FN(sub)(D, A);
  REAL Ds[9], F[9], dF[9], MT[9], g[9], gt[9];
  FN(mat_cols)(Ds, p1, p2, p3);
  FN(mat_mul)(F, Ds, M);
  V3 F0 = FN(row)(F, 0), F1 = FN(row)(F, 1), F2 = FN(row)(F, 2);
  V3 dF1 = FN(cross)(F1, F2), dF2 = FN(cross)(F2, F0), dF3 = FN(cross)(F0, F1);
  FN(mat_cols)(dF, dF1, dF2, dF3);
  FN(mat_T)(MT, M);
  FN(mat_mul)(g, dF, MT);
  FN(mat_T)(gt, g);
  o.gH[1] = FN(row)(gt, 0); o.gH[2] = FN(row)(gt, 1); o.gH[3] = FN(row)(gt, 2);
  o.gH[0] = FN(sub)(FN(sub)(FN(neg)(o.gH[1]), o.gH[2]), o.gH[3]);
  REAL volume = FN(dot)(F0, FN(cross)(F1, F2));
  REAL invV = R(1.0) / volC, invD = R(1.0) / devC;
  o.CH = volume - R(1.0) - invD / invV;
  REAL l0 = FN(length)(F0), l1 = FN(length)(F1), l2 = FN(length)(F2);
  REAL rs = FN(sqrt_)(l0 * l0 + l1 * l1 + l2 * l2);
  REAL rsi = R(1.0) / rs;
  o.CD = rs;
  FN(mat_mul)(g, F, MT);
  for (int i = 0; i < 9; ++i) g[i] = g[i] * rsi;
  FN(mat_T)(gt, g);
  o.gD[1] = FN(row)(gt, 0); o.gD[2] = FN(row)(gt, 1); o.gD[3] = FN(row)(gt, 2);
  o.gD[0] = FN(sub)(FN(sub)(FN(neg)(o.gD[1]), o.gD[2]), o.gD[3]);
  return o;
}

static inline REAL FN(nh_w)(const V3 g[4], const REAL m[4]) {
  REAL l, w = R(0);
  for (int i = 0; i < 4; ++i) { l = FN(length)(g[i]); w += l * l * m[i]; }
  return w;
}

/* Debug export of nh_terms: out = [C_H, gH (4x3), C_D, gD (4x3)]. */
void FN(oracle_nh_terms)(const REAL* pos, const REAL* M, REAL volC, REAL devC, REAL* out) {
  FN(nh_terms_t) o = FN(nh_terms)(FN(ld)(pos, 0), FN(ld)(pos, 1), FN(ld)(pos, 2), FN(ld)(pos, 3), M, volC, devC);
  out[0] = o.CH;
  for (int i = 0; i < 4; ++i) FN(st)(out + 1, i, o.gH[i]);
  out[13] = o.CD;
  for (int i = 0; i < 4; ++i) FN(st)(out + 14, i, o.gD[i]);
}

/* neoHookeanConstraints (integrator.py:578-698).  lambda is read and never
 * stored, and no alpha*lambda term enters dlambda, exactly as the reference. */
void FN(oracle_neohookean)(long nT, const REAL* pred, REAL* dP, int* cnt, const int* tet,
                           const REAL* restVol, const REAL* invMass, const REAL* invRest,
                           REAL volC, REAL devC, REAL dT) {
  for (long tid = 0; tid < nT; ++tid) {
    const int* t = tet + 4 * tid;
    REAL rv = restVol[tid];
    REAL m[4] = {invMass[t[0]], invMass[t[1]], invMass[t[2]], invMass[t[3]]};
    FN(nh_terms_t) o = FN(nh_terms)(FN(ld)(pred, t[0]), FN(ld)(pred, t[1]), FN(ld)(pred, t[2]),
                                    FN(ld)(pred, t[3]), invRest + 9 * tid, volC, devC);
    REAL w = FN(nh_w)(o.gH, m);
    if (w < FEPS) continue;
    REAL alpha = volC / dT / dT / rv;
    REAL dl = -o.CH / (w + alpha);
    V3 d[4];
    for (int i = 0; i < 4; ++i) d[i] = FN(mul)(FN(mul)(o.gH[i], dl), m[i]);
    w = FN(nh_w)(o.gD, m);
    if (w < FEPS) continue;
    alpha = devC / dT / dT / rv;
    dl = -o.CD / (w + alpha);
    for (int i = 0; i < 4; ++i) d[i] = FN(add)(d[i], FN(mul)(FN(mul)(o.gD[i], dl), m[i]));
    FN(scatter4)(dP, cnt, t, d);
  }
}

/* applyConstraints (integrator.py:946-960) */
void FN(oracle_apply)(long n, REAL* pred, REAL* dP, int* cnt) {
  for (long tid = 0; tid < n; ++tid) {
    if (cnt[tid] > 0) {
      V3 d = FN(ld)(dP, tid);
      REAL N = (REAL)cnt[tid];
      FN(st)(pred, tid, FN(add)(FN(ld)(pred, tid), FN(v3)(d.x / N, d.y / N, d.z / N)));
    }
    FN(st)(dP, tid, FN(v3)(R(0), R(0), R(0)));
    cnt[tid] = 0;
  }
}

/* dragConstraintsDO (integrator.py:412-442): pull toward tip[env*4+2] */
void FN(oracle_drag)(long n, REAL* pred, const REAL* drag, const int* vertexToEnv, const REAL* tips,
                     REAL ks) {
  for (long tid = 0; tid < n; ++tid) {
    REAL dr = drag[tid];
    if (dr == R(0)) continue;
    int env = vertexToEnv[tid];
    V3 p = FN(ld)(pred, tid);
    V3 tip = FN(ld)(tips, (long)env * 4 + 2);
    V3 dir = FN(sub)(p, tip);
    REAL len = FN(length)(dir);
    if (len < FEPS) continue;
    V3 e = FN(divs)(FN(mul)(FN(mul)(FN(mul)(FN(normalize)(dir), len), dr), ks), R(2.0));
    FN(st)(pred, tid, FN(sub)(p, e));
  }
}

/* bounds (integrator.py:12-21) */
void FN(oracle_bounds)(long n, REAL* pred, REAL ground) {
  for (long tid = 0; tid < n; ++tid)
    if (pred[3 * tid + 1] < ground) pred[3 * tid + 1] = ground;
}

/* PBDStep (integrator.py:23-38) */
void FN(oracle_pbd_step)(long n, REAL* vertex, const REAL* pred, REAL* vel, REAL dT) {
  REAL inv = R(1.0) / dT;
  for (long tid = 0; tid < n; ++tid) {
    V3 x = FN(ld)(vertex, tid), p = FN(ld)(pred, tid);
    FN(st)(vel, tid, FN(mul)(FN(sub)(p, x), inv));
    FN(st)(vertex, tid, p);
  }
}

/* updateClampsDragKernel (modelDO.py:247-279): radius selection around tip[env*4+3] */
void FN(oracle_grasp_select)(long n, const REAL* pos, const int* vertexToEnv, const REAL* tips,
                             REAL* drag, REAL radius, const int* state, const int* update) {
  for (long tid = 0; tid < n; ++tid) {
    int env = vertexToEnv[tid];
    if (update[env * 3] == 0) continue;
    if (state[env * 3] == 1) {
      REAL len = FN(length)(FN(sub)(FN(ld)(pos, tid), FN(ld)(tips, (long)env * 4 + 3)));
      drag[tid] = len < radius ? R(1.0) : R(0.0);
    } else {
      drag[tid] = R(0.0);
    }
  }
}

/* closestPointOnEdge (integrator.py:1521-1534) */
static inline V3 FN(closest_on_edge)(V3 p, V3 ea, V3 eb) {
  V3 d = FN(sub)(eb, ea);
  REAL t;
  if (FN(length)(d) < FEPS) {
    t = R(0.5);
  } else {
    REAL d2 = FN(length)(d) * FN(length)(d);
    t = FN(dot)(d, FN(sub)(p, ea)) / d2;
    if (t < R(0)) t = R(0);
    else if (t > R(1)) t = R(1);
  }
  return FN(add)(ea, FN(mul)(d, t));
}

/* triangle_closest_point_barycentric (integrator.py:1536-1582) */
static inline V3 FN(closest_bary)(V3 a, V3 b, V3 c, V3 p) {
  V3 ab = FN(sub)(b, a), ac = FN(sub)(c, a), ap = FN(sub)(p, a);
  REAL d1 = FN(dot)(ab, ap), d2 = FN(dot)(ac, ap);
  if (d1 <= R(0) && d2 <= R(0)) return FN(v3)(R(1), R(0), R(0));
  V3 bp = FN(sub)(p, b);
  REAL d3 = FN(dot)(ab, bp), d4 = FN(dot)(ac, bp);
  if (d3 >= R(0) && d4 <= d3) return FN(v3)(R(0), R(1), R(0));
  REAL vc = d1 * d4 - d3 * d2;
  REAL v = d1 / (d1 - d3);
  if (vc <= R(0) && d1 >= R(0) && d3 <= R(0)) return FN(v3)(R(1) - v, v, R(0));
  V3 cp = FN(sub)(p, c);
  REAL d5 = FN(dot)(ab, cp), d6 = FN(dot)(ac, cp);
  if (d6 >= R(0) && d5 <= d6) return FN(v3)(R(0), R(0), R(1));
  REAL vb = d5 * d2 - d1 * d6;
  REAL w = d2 / (d2 - d6);
  if (vb <= R(0) && d2 >= R(0) && d6 <= R(0)) return FN(v3)(R(1) - w, R(0), w);
  REAL va = d3 * d6 - d5 * d4;
  w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
  if (va <= R(0) && (d4 - d3) >= R(0) && (d5 - d6) >= R(0)) return FN(v3)(R(0), w, R(1) - w);
  REAL denom = R(1) / (va + vb + vc);
  v = vb * denom;
  w = vc * denom;
  return FN(v3)(R(1) - v - w, v, w);
}

/* checkSdfCapsuleCollision (integrator.py:1303-1355): the response of each
 * triangle vertex to one capsule (r[0..2]); returns 0 when there is none.
 * As in the reference, the barycentric coordinates of the line/plane
 * intersection (not the point) are what closestPointOnEdge receives. */
static inline int FN(capsule_response)(V3 tip, V3 base, REAL radius, V3 triNormal, V3 triCentre, V3 pA, V3 pB,
                                       V3 pC, REAL kS, V3 r[3]) {
  V3 cn = FN(normalize)(FN(sub)(tip, base));
  V3 off = FN(mul)(cn, radius);
  V3 posA = FN(add)(base, off), posB = FN(sub)(tip, off);
  REAL denom = FN(dot)(triNormal, cn);
  if (FN(absr)(denom) < FEPS) return 0;
  REAL t = FN(dot)(triNormal, FN(sub)(triCentre, base)) / denom;
  V3 lpi = FN(add)(base, FN(mul)(cn, t));
  V3 refp = FN(closest_bary)(pA, pB, pC, lpi);
  V3 sc = FN(closest_on_edge)(refp, posA, posB);
  V3 bary = FN(closest_bary)(pA, pB, pC, sc);
  V3 contact = FN(add)(FN(add)(FN(mul)(pA, bary.x), FN(mul)(pB, bary.y)), FN(mul)(pC, bary.z));
  V3 sv = FN(sub)(sc, contact);
  REAL distance = FN(length)(sv);
  REAL pen = radius / FN(length)(sv) - R(1.0);
  V3 pv = FN(v3)(pen * sv.x, pen * sv.y, pen * sv.z);
  REAL plen = FN(length)(pv);
  V3 pn = FN(normalize)(pv);
  REAL s = bary.x * bary.x + bary.y * bary.y + bary.z * bary.z;
  if (s < FEPS) return 0;
  V3 resp = FN(mul)(FN(mul)(pn, plen / s), kS);
  if (distance > radius) return 0;
  r[0] = FN(mul)(resp, bary.x);
  r[1] = FN(mul)(resp, bary.y);
  r[2] = FN(mul)(resp, bary.z);
  return 1;
}

/* collideLaparoscopeDO (integrator.py:981-1050): every surface triangle
 * against the first three capsules of its env's instrument (base, tip,
 * radius at env*4 + i); inCollision[env] counts the triangles with a
 * response.  The responses are computed from the positions at the start of
 * the pass and then subtracted from the vertices with inverse mass > 0 (the
 * reference subtracts with unordered atomics on the GPU, so no triangle sees
 * another's update in a defined order; this is that order-free form).
 * tri: F x 3 global vertex ids; acc: nV x 3 scratch, left zero. */
void FN(oracle_collide)(long nF, REAL* pred, const int* tri, const int* triToEnv, const REAL* invMass,
                        const REAL* capsBase, const REAL* capsTip, const REAL* capsRadius, int* inCollision,
                        REAL kS, long nV, REAL* acc) {
  for (long v = 0; v < nV; ++v) FN(st)(acc, v, FN(v3)(R(0), R(0), R(0)));
  for (long tid = 0; tid < nF; ++tid) {
    int env = triToEnv[tid];
    int ia = tri[3 * tid], ib = tri[3 * tid + 1], ic = tri[3 * tid + 2];
    V3 pA = FN(ld)(pred, ia), pB = FN(ld)(pred, ib), pC = FN(ld)(pred, ic);
    V3 centre = FN(mul)(FN(add)(FN(add)(pA, pB), pC), R(0.3333));
    V3 n = FN(normalize)(FN(cross)(FN(sub)(pB, pA), FN(sub)(pC, pA)));
    V3 tA = FN(v3)(R(0), R(0), R(0)), tB = tA, tC = tA;
    REAL len = R(0);
    for (int i = 0; i < 3; ++i) {
      V3 r[3] = {FN(v3)(R(0), R(0), R(0)), FN(v3)(R(0), R(0), R(0)), FN(v3)(R(0), R(0), R(0))};
      FN(capsule_response)(FN(ld)(capsTip, env * 4 + i), FN(ld)(capsBase, env * 4 + i), capsRadius[env * 4 + i], n,
                           centre, pA, pB, pC, kS, r);
      tA = FN(add)(tA, r[0]);
      tB = FN(add)(tB, r[1]);
      tC = FN(add)(tC, r[2]);
      len = len + FN(length)(r[0]) + FN(length)(r[1]) + FN(length)(r[2]);
    }
    if (len > R(0)) inCollision[env] += 1;
    if (invMass[ia] > R(0)) FN(st)(acc, ia, FN(add)(FN(ld)(acc, ia), tA));
    if (invMass[ib] > R(0)) FN(st)(acc, ib, FN(add)(FN(ld)(acc, ib), tB));
    if (invMass[ic] > R(0)) FN(st)(acc, ic, FN(add)(FN(ld)(acc, ic), tC));
  }
  for (long v = 0; v < nV; ++v) {
    FN(st)(pred, v, FN(sub)(FN(ld)(pred, v), FN(ld)(acc, v)));
    FN(st)(acc, v, FN(v3)(R(0), R(0), R(0)));
  }
}

/* Constraint set of one step (the arguments SimModelDO hands to the kernels). */
typedef struct {
  long nV, nE, nT;
  int substeps, iterations;
  REAL dt;            /* substep dt = 1/(fps*substeps) */
  REAL gravity[3], damping, ground;
  int dist_mode;      /* 0 off, 1 PBD (kS), 2 XPBD (alpha~) */
  int vol_mode;       /* 0 off, 1 PBD (kS), 2 XPBD (alpha~) */
  int nh_on;          /* 1: neoHookeanConstraints */
  int lambda_legacy;  /* 1: zero lambdas only when iterations > 1 (legacy stepper) */
  REAL dist_k, vol_k; /* kS or alpha~ depending on mode */
  REAL nh_vol_c, nh_dev_c;
  REAL drag_ks;
  int collide;        /* 1: collideLaparoscopeDO after the grasp */
  long nF;            /* surface triangles (collision) */
  REAL contact_k;     /* globalKsContact */
} FN(oracle_step_cfg);

/* SimIntegratorDO.stepModel (integrator.py:1640-1744), one frame.  Lambdas are
 * zeroed at the start of each substep (or as the legacy stepper does,
 * integrator.py:2154-2173). */
void FN(oracle_step_frame)(const FN(oracle_step_cfg)* c, REAL* vertex, REAL* pred, REAL* vel,
                           const REAL* invMass, REAL* dP, int* cnt, const int* edge,
                           const REAL* edgeRest, const int* tet, const REAL* restVol,
                           const REAL* invRest, REAL* lamE, REAL* lamT, const REAL* drag,
                           const int* vertexToEnv, const REAL* tips, const int* tri, const int* triToEnv,
                           const REAL* capsBase, const REAL* capsRadius, int* inCollision) {
  for (int s = 0; s < c->substeps; ++s) {
    FN(oracle_gravity)(c->nV, vertex, pred, vel, invMass, c->gravity, c->damping, c->dt);
    if (!c->lambda_legacy || c->iterations > 1) {
      for (long i = 0; i < c->nE; ++i) lamE[i] = R(0);
      for (long i = 0; i < c->nT; ++i) lamT[i] = R(0);
    }
    for (int it = 0; it < c->iterations; ++it) {
      if (c->vol_mode == 1) FN(oracle_volume_pbd)(c->nT, pred, dP, cnt, tet, restVol, invMass, 0, c->vol_k);
      if (c->vol_mode == 2) FN(oracle_volume_xpbd)(c->nT, pred, dP, lamT, cnt, tet, restVol, invMass, 0, c->vol_k);
      if (c->dist_mode == 1) FN(oracle_distance_pbd)(c->nE, pred, dP, cnt, edge, edgeRest, invMass, 0, c->dist_k);
      if (c->dist_mode == 2) FN(oracle_distance_xpbd)(c->nE, pred, dP, lamE, cnt, edge, edgeRest, invMass, 0, c->dist_k);
      if (c->nh_on) FN(oracle_neohookean)(c->nT, pred, dP, cnt, tet, restVol, invMass, invRest, c->nh_vol_c, c->nh_dev_c, c->dt);
      FN(oracle_apply)(c->nV, pred, dP, cnt);
      if (drag) FN(oracle_drag)(c->nV, pred, drag, vertexToEnv, tips, c->drag_ks);
      if (c->collide)
        FN(oracle_collide)(c->nF, pred, tri, triToEnv, invMass, capsBase, tips, capsRadius, inCollision,
                           c->contact_k, c->nV, dP);
    }
    FN(oracle_bounds)(c->nV, pred, c->ground);
    FN(oracle_pbd_step)(c->nV, vertex, pred, vel, c->dt);
  }
}

/* refitBVHLevel + refitBVHLoop (bvh.py:30-104, 419-440) for one environment
 * whose faces index its own vertex array (visFaceShift = 0). */
void FN(oracle_refit)(int numNodes, int maxLevel, const int* leftNode, const int* firstTri,
                      const int* triCount, const int* level, const int* triMap, const int* faces,
                      const REAL* vertex, REAL* aabbMin, REAL* aabbMax) {
  for (int L = maxLevel; L >= 0; --L) {
    for (int node = 0; node < numNodes; ++node) {
      if (level[node] != L) continue;
      int left = leftNode[node];
      if (left < 0) {
        V3 mn = FN(v3)(R(1e6), R(1e6), R(1e6)), mx = FN(v3)(R(-1e6), R(-1e6), R(-1e6));
        for (int k = 0; k < triCount[node]; ++k) {
          int tri = triMap[firstTri[node] + k];
          V3 a = FN(ld)(vertex, faces[3 * tri]), b = FN(ld)(vertex, faces[3 * tri + 1]),
             c = FN(ld)(vertex, faces[3 * tri + 2]);
          mn = FN(vmin)(a, FN(vmin)(b, FN(vmin)(c, mn)));
          mx = FN(vmax)(a, FN(vmax)(b, FN(vmax)(c, mx)));
        }
        FN(st)(aabbMin, node, mn); FN(st)(aabbMax, node, mx);
      } else {
        int right = left + 1;
        FN(st)(aabbMin, node, FN(vmin)(FN(ld)(aabbMin, left), FN(ld)(aabbMin, right)));
        FN(st)(aabbMax, node, FN(vmax)(FN(ld)(aabbMax, left), FN(ld)(aabbMax, right)));
      }
    }
  }
}

/* intersectAABB (render.py:588-609): returns hit and writes tmin */
static inline int FN(isect_aabb)(V3 o, V3 d, V3 lo, V3 hi, REAL* tmin_out) {
  REAL t1 = (lo.x - o.x) / d.x, t2 = (hi.x - o.x) / d.x;
  REAL t3 = (lo.y - o.y) / d.y, t4 = (hi.y - o.y) / d.y;
  REAL t5 = (lo.z - o.z) / d.z, t6 = (hi.z - o.z) / d.z;
  REAL tmin = FN(fmax_)(FN(fmax_)(FN(fmin_)(t1, t2), FN(fmin_)(t3, t4)), FN(fmin_)(t5, t6));
  REAL tmax = FN(fmin_)(FN(fmin_)(FN(fmax_)(t1, t2), FN(fmax_)(t3, t4)), FN(fmax_)(t5, t6));
  if (tmax >= R(0) && tmax >= tmin) { *tmin_out = tmin; return 1; }
  *tmin_out = R(0);
  return 0;
}

/* MTTriRayIntersection (render.py:661-697): returns t, or -1 on a miss */
static inline REAL FN(mt)(V3 a, V3 b, V3 c, V3 o, V3 d) {
  V3 e1 = FN(sub)(b, a), e2 = FN(sub)(c, a);
  V3 pv = FN(cross)(d, e2);
  REAL det = FN(dot)(e1, pv);
  if (FN(absr)(det) < R(0.0001)) return R(-1.0);
  REAL inv = R(1.0) / det;
  V3 tv = FN(sub)(o, a);
  REAL u = FN(dot)(tv, pv) * inv;
  if (u < R(0) || u > R(1)) return R(-1.0);
  V3 qv = FN(cross)(tv, e1);
  REAL v = FN(dot)(d, qv) * inv;
  if (v < R(0) || u + v > R(1)) return R(-1.0);
  return FN(dot)(e2, qv) * inv;
}

/* meshQueryRayManyEnvs (render.py:411-498) for one ray.  Returns the hit face
 * (an index into faces) or -1, and its t.  *max_stack reports the deepest
 * stack use (the reference's stack holds 32 entries). */
int FN(oracle_ray)(V3 o, V3 d, const REAL* aabbMin, const REAL* aabbMax, const int* leftNode,
                   const int* firstTri, const int* triCount, const int* triMap, const int* faces,
                   const REAL* vertex, REAL* t_out, int* max_stack) {
  int stack[256];
  int count = 1, hit = -1;
  stack[0] = 0;
  REAL tMin = R(1e6);
  while (count) {
    int node = stack[count - 1];
    count -= 1;
    REAL tb;
    int h = FN(isect_aabb)(o, d, FN(ld)(aabbMin, node), FN(ld)(aabbMax, node), &tb);
    if (h && tb < tMin) {
      int left = leftNode[node];
      if (left == -1) {
        for (int k = 0; k < triCount[node]; ++k) {
          int tri = triMap[firstTri[node] + k];
          V3 a = FN(ld)(vertex, faces[3 * tri]), b = FN(ld)(vertex, faces[3 * tri + 1]),
             c = FN(ld)(vertex, faces[3 * tri + 2]);
          REAL t = FN(mt)(a, b, c, o, d);
          if (t < tMin && t > R(0)) { tMin = t; hit = tri; }
        }
      } else {
        if (count + 2 > 256) return -2;
        stack[count] = left;
        stack[count + 1] = left + 1;
        count += 2;
        if (count > *max_stack) *max_stack = count;
      }
    }
  }
  if (tMin >= R(1e6)) { *t_out = R(0); return -1; }
  *t_out = tMin;
  return hit;
}

/* drawSimBVHKernelManyEnvs (render.py:276-409) depth mode (color = t/50 clamped
 * to 1, miss = 0) for one environment, plus downsample_kernel (render.py:699-724)
 * with a 2x2 box.  dirs holds one direction per ray of the (2H x 2W) ray grid in
 * row-major ray order; hit/tray receive per-ray face ids and t (may be NULL). */
void FN(oracle_render_depth)(int H, int W, const REAL* origin, const REAL* dirs,
                             const REAL* aabbMin, const REAL* aabbMax, const int* leftNode,
                             const int* firstTri, const int* triCount, const int* triMap,
                             const int* faces, const REAL* vertex, REAL* depth, int* hit, REAL* tray,
                             int* max_stack) {
  int RW = 2 * W, RH = 2 * H;
  V3 o = FN(v3)(origin[0], origin[1], origin[2]);
  REAL* ray = (REAL*)malloc(sizeof(REAL) * (size_t)RW * RH);
  for (int r = 0; r < RW * RH; ++r) {
    REAL t;
    int h = FN(oracle_ray)(o, FN(ld)(dirs, r), aabbMin, aabbMax, leftNode, firstTri, triCount, triMap,
                           faces, vertex, &t, max_stack);
    REAL col = R(0);
    if (h > -1) {
      col = t / R(50.0) * R(1.0);
      if (col > R(1.0)) col = R(1.0);
    }
    ray[r] = col;
    if (hit) hit[r] = h;
    if (tray) tray[r] = t;
  }
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      int start = py * 2 * RW + px * 2;
      REAL c = R(0);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) c = c + ray[start + i * RW + j];
      depth[py * W + px] = c / R(4.0);
    }
  free(ray);
}


/* zeroVec3Array + vertex_normal_sum_kernel + normalize_kernel (render.py:81-121):
 * area-weighted vertex normals, summed in serial face (tid) order.
 * faces: F x 3, vertex / normals: nv x 3. */
void FN(oracle_vertex_normals)(long nf, const int* faces, long nv, const REAL* vertex, REAL* normals) {
  for (long v = 0; v < nv; ++v) FN(st)(normals, v, FN(v3)(R(0), R(0), R(0)));
  for (long tid = 0; tid < nf; ++tid) {
    int i = faces[3 * tid], j = faces[3 * tid + 1], k = faces[3 * tid + 2];
    V3 a = FN(ld)(vertex, i), b = FN(ld)(vertex, j), c = FN(ld)(vertex, k);
    V3 n = FN(cross)(FN(sub)(b, a), FN(sub)(c, a));
    FN(st)(normals, i, FN(add)(FN(ld)(normals, i), n));
    FN(st)(normals, j, FN(add)(FN(ld)(normals, j), n));
    FN(st)(normals, k, FN(add)(FN(ld)(normals, k), n));
  }
  for (long v = 0; v < nv; ++v) FN(st)(normals, v, FN(normalize)(FN(ld)(normals, v)));
}

/* float -> int as Warp's int() on the device: truncation toward zero;
 * NaN gives 0 and values beyond the int range saturate. */
static inline int FN(trunc_i)(REAL x) {
  if (x != x) return 0;
  if (x >= R(2147483647.0)) return 2147483647;
  if (x <= R(-2147483648.0)) return -2147483647 - 1;
  return (int)x;
}

/* One texel of the concatenated texture table (RGB triples, n_texels of
 * them); an index outside the table reads as zero (the reference reads past
 * a texture's last column, where its weight is zero). */
static inline V3 FN(texel)(const REAL* texels, long n_texels, long i) {
  if (i < 0 || i >= n_texels) return FN(v3)(R(0), R(0), R(0));
  return FN(ld)(texels, i);
}

/* texture_interpolation (render.py:124-148): bilinear lookup of a square
 * texture of side `size` starting at texel `shift`, with the reference's
 * indexing y + x * width (x along the first texture coordinate) and no clamp
 * of x0 + 1 / y0 + 1. */
V3 FN(oracle_texture_interp)(REAL s, REAL t, const REAL* texels, long n_texels, long shift, int size) {
  REAL tx = s * (REAL)(size - 1), ty = (R(1.0) - t) * (REAL)(size - 1);
  int x0 = FN(trunc_i)(tx), y0 = FN(trunc_i)(ty);
  int x1 = x0 + 1, y1 = y0 + 1;
  REAL ax = tx - (REAL)x0, ay = ty - (REAL)y0;
  V3 c00 = FN(texel)(texels, n_texels, shift + y0 + (long)x0 * size);
  V3 c10 = FN(texel)(texels, n_texels, shift + y0 + (long)x1 * size);
  V3 c01 = FN(texel)(texels, n_texels, shift + y1 + (long)x0 * size);
  V3 c11 = FN(texel)(texels, n_texels, shift + y1 + (long)x1 * size);
  V3 lower = FN(add)(FN(mul)(c00, R(1.0) - ax), FN(mul)(c10, ax));
  V3 upper = FN(add)(FN(mul)(c01, R(1.0) - ax), FN(mul)(c11, ax));
  return FN(add)(FN(mul)(lower, R(1.0) - ay), FN(mul)(upper, ay));
}

/* Entry point of the texture lookup for tests: out = RGB at (s, t). */
void FN(oracle_texture_lookup)(REAL s, REAL t, const REAL* texels, long n_texels, long shift, int size, REAL* out) {
  FN(st)(out, 0, FN(oracle_texture_interp)(s, t, texels, n_texels, shift, size));
}

/* drawSimBVHKernelManyEnvs (render.py:276-409) colour modes for one
 * environment, plus downsample_kernel (render.py:699-724) per channel.
 *   mode 0 grayscale; 1 texture (render.py:356-380): the hit triangle's
 *   texture coordinates u tex_a + v tex_b + w tex_c (tri_uv: 6 per face id),
 *   looked up in texture tri_tex[face] (tex_info: first texel, side per
 *   texture), or, for a triangle without a texture (tri_tex < 0), the mean
 *   of its vertex colours; 3 vertex colour; 5 per-triangle colour (tcolor,
 *   this build's seeded colours); each times the Lambert term
 *   max(0, intensity * dot(normal, light)) of the single light whose
 *   direction is the raw light position (render.py:387-393); 2 normal map.
 * Barycentrics from cross-product lengths, normal = u na + v nb + w nc
 * (render.py:329-350).  Each channel is clamped to 1.  rgb: H x W x 3. */
void FN(oracle_render_rgb)(int H, int W, int mode, const REAL* origin, const REAL* dirs,
                           const REAL* aabbMin, const REAL* aabbMax, const int* leftNode,
                           const int* firstTri, const int* triCount, const int* triMap,
                           const int* faces, const REAL* vertex, const REAL* normals, const REAL* vcolor,
                           const REAL* tcolor, const REAL* tri_uv, const int* tri_tex, const REAL* texels,
                           long n_texels, const long* tex_info, const REAL* light, REAL intensity, REAL* rgb) {
  int RW = 2 * W, RH = 2 * H, ms = 0;
  V3 o = FN(v3)(origin[0], origin[1], origin[2]);
  V3 L = FN(v3)(light[0], light[1], light[2]);
  REAL* ray = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)RW * RH);
  for (int r = 0; r < RW * RH; ++r) {
    REAL t;
    V3 d = FN(ld)(dirs, r);
    int h = FN(oracle_ray)(o, d, aabbMin, aabbMax, leftNode, firstTri, triCount, triMap, faces, vertex, &t, &ms);
    V3 color = FN(v3)(R(0), R(0), R(0));
    if (h > -1) {
      int i = faces[3 * h], j = faces[3 * h + 1], k = faces[3 * h + 2];
      V3 a = FN(ld)(vertex, i), b = FN(ld)(vertex, j), c = FN(ld)(vertex, k);
      V3 p = FN(add)(o, FN(mul)(d, t));
      REAL area = FN(length)(FN(cross)(FN(sub)(b, a), FN(sub)(c, a)));
      REAL w = FN(length)(FN(cross)(FN(sub)(b, a), FN(sub)(p, a))) / area;
      REAL v = FN(length)(FN(cross)(FN(sub)(p, a), FN(sub)(c, a))) / area;
      REAL u = R(1.0) - w - v;
      V3 normal = FN(add)(FN(add)(FN(mul)(FN(ld)(normals, i), u), FN(mul)(FN(ld)(normals, j), v)),
                          FN(mul)(FN(ld)(normals, k), w));
      if (mode == 2) {
        color = FN(add)(FN(mul)(normal, R(0.5)), FN(v3)(R(0.5), R(0.5), R(0.5)));
      } else {
        if (mode == 0) {
          color = FN(v3)(R(1), R(1), R(1));
        } else if (mode == 5) {
          color = FN(ld)(tcolor, h);
        } else if (mode == 1 && tri_tex[h] >= 0) {
          const REAL* uv = tri_uv + 6 * (long)h;
          REAL s = (u * uv[0] + v * uv[2]) + w * uv[4];
          REAL tt = (u * uv[1] + v * uv[3]) + w * uv[5];
          const long* ti = tex_info + 2 * (long)tri_tex[h];
          color = FN(oracle_texture_interp)(s, tt, texels, n_texels, ti[0], (int)ti[1]);
        } else {  /* vertex colour, and texture mode's fallback */
          color = FN(divs)(FN(add)(FN(add)(FN(ld)(vcolor, i), FN(ld)(vcolor, j)), FN(ld)(vcolor, k)), R(3.0));
        }
        REAL val = intensity * FN(dot)(normal, L);
        if (val < R(0)) val = R(0);
        REAL lambert = R(0.0) + val;
        color = FN(mul)(color, lambert);
      }
      if (color.x > R(1)) color.x = R(1);
      if (color.y > R(1)) color.y = R(1);
      if (color.z > R(1)) color.z = R(1);
    }
    FN(st)(ray, r, color);
  }
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      int start = py * 2 * RW + px * 2;
      V3 c = FN(v3)(R(0), R(0), R(0));
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) c = FN(add)(c, FN(ld)(ray, start + i * RW + j));
      FN(st)(rgb, (long)py * W + px, FN(divs)(c, R(4.0)));
    }
  free(ray);
}

/* ---- instrument kinematics (modelDO.py:55-289, 1331-1404, 1570-1587) ----
 * Transforms are 4x4, stored transposed as the reference stores them
 * (xf[(env*7+k)*16 + 4*row + col] is the transposed matrix).  A4 below is a
 * column-vector matrix a[4*i + j] = row i, column j.  Products sum k = 0..3
 * left to right.  Warp's mat44 inverse, quat_from_axis_angle, quaternion to
 * matrix and asin are restated from their definitions; their exact operation
 * order in warp-lang 0.10.1 is parity unpinned (tests pin the geometry). */
#define A4 FN(mat4_t)
typedef struct { REAL a[16]; } A4;
static inline A4 FN(m4_load)(const REAL* xf, int id) {
  A4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.a[4 * i + j] = xf[id * 16 + 4 * j + i];
  return r;
}
static inline void FN(m4_store)(REAL* xf, int id, A4 m) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) xf[id * 16 + 4 * j + i] = m.a[4 * i + j];
}
static inline A4 FN(m4_T)(A4 m) {
  A4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.a[4 * i + j] = m.a[4 * j + i];
  return r;
}
static inline A4 FN(m4_mul)(A4 x, A4 y) {
  A4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      REAL s = R(0);
      for (int k = 0; k < 4; ++k) s = s + x.a[4 * i + k] * y.a[4 * k + j];
      r.a[4 * i + j] = s;
    }
  return r;
}
/* adjugate / determinant with the 2x2 minors of rows (0,1) and (2,3) */
static inline A4 FN(m4_inv)(A4 x) {
#define E(i, j) x.a[4 * (i) + (j)]
  REAL s0 = E(0, 0) * E(1, 1) - E(1, 0) * E(0, 1), s1 = E(0, 0) * E(1, 2) - E(1, 0) * E(0, 2);
  REAL s2 = E(0, 0) * E(1, 3) - E(1, 0) * E(0, 3), s3 = E(0, 1) * E(1, 2) - E(1, 1) * E(0, 2);
  REAL s4 = E(0, 1) * E(1, 3) - E(1, 1) * E(0, 3), s5 = E(0, 2) * E(1, 3) - E(1, 2) * E(0, 3);
  REAL c5 = E(2, 2) * E(3, 3) - E(3, 2) * E(2, 3), c4 = E(2, 1) * E(3, 3) - E(3, 1) * E(2, 3);
  REAL c3 = E(2, 1) * E(3, 2) - E(3, 1) * E(2, 2), c2 = E(2, 0) * E(3, 3) - E(3, 0) * E(2, 3);
  REAL c1 = E(2, 0) * E(3, 2) - E(3, 0) * E(2, 2), c0 = E(2, 0) * E(3, 1) - E(3, 0) * E(2, 1);
  REAL det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  A4 r;
  if (det == R(0)) {
    for (int i = 0; i < 16; ++i) r.a[i] = R(0);
    return r;
  }
  REAL id = R(1) / det;
  r.a[0] = (E(1, 1) * c5 - E(1, 2) * c4 + E(1, 3) * c3) * id;
  r.a[1] = (-E(0, 1) * c5 + E(0, 2) * c4 - E(0, 3) * c3) * id;
  r.a[2] = (E(3, 1) * s5 - E(3, 2) * s4 + E(3, 3) * s3) * id;
  r.a[3] = (-E(2, 1) * s5 + E(2, 2) * s4 - E(2, 3) * s3) * id;
  r.a[4] = (-E(1, 0) * c5 + E(1, 2) * c2 - E(1, 3) * c1) * id;
  r.a[5] = (E(0, 0) * c5 - E(0, 2) * c2 + E(0, 3) * c1) * id;
  r.a[6] = (-E(3, 0) * s5 + E(3, 2) * s2 - E(3, 3) * s1) * id;
  r.a[7] = (E(2, 0) * s5 - E(2, 2) * s2 + E(2, 3) * s1) * id;
  r.a[8] = (E(1, 0) * c4 - E(1, 1) * c2 + E(1, 3) * c0) * id;
  r.a[9] = (-E(0, 0) * c4 + E(0, 1) * c2 - E(0, 3) * c0) * id;
  r.a[10] = (E(3, 0) * s4 - E(3, 1) * s2 + E(3, 3) * s0) * id;
  r.a[11] = (-E(2, 0) * s4 + E(2, 1) * s2 - E(2, 3) * s0) * id;
  r.a[12] = (-E(1, 0) * c3 + E(1, 1) * c1 - E(1, 2) * c0) * id;
  r.a[13] = (E(0, 0) * c3 - E(0, 1) * c1 + E(0, 2) * c0) * id;
  r.a[14] = (-E(3, 0) * s3 + E(3, 1) * s1 - E(3, 2) * s0) * id;
  r.a[15] = (E(2, 0) * s3 - E(2, 1) * s1 + E(2, 2) * s0) * id;
#undef E
  return r;
}
static inline V3 FN(m4_point)(A4 m, V3 p) {
  REAL o[3];
  for (int i = 0; i < 3; ++i) o[i] = ((m.a[4 * i] * p.x + m.a[4 * i + 1] * p.y) + m.a[4 * i + 2] * p.z) + m.a[4 * i + 3];
  return FN(v3)(o[0], o[1], o[2]);
}
/* quat_from_axis_angle: (axis sin(a/2), cos(a/2)); q[3] = w */
static inline void FN(quat)(V3 axis, REAL angle, REAL q[4]) {
  REAL h = angle * R(0.5), s = (REAL)SIN(h);
  q[0] = axis.x * s; q[1] = axis.y * s; q[2] = axis.z * s; q[3] = (REAL)COS(h);
}
/* quat_rotate: x (2w^2 - 1) + (v x x) 2w + v 2(v . x) */
static inline V3 FN(qrot)(const REAL q[4], V3 x) {
  V3 v = FN(v3)(q[0], q[1], q[2]);
  REAL c = R(2) * q[3] * q[3] - R(1), d = R(2) * FN(dot)(v, x);
  return FN(add)(FN(add)(FN(mul)(x, c), FN(mul)(FN(cross)(v, x), R(2) * q[3])), FN(mul)(v, d));
}
/* wp.mat44(pos, rot, scale=1): rotation columns = rotated axes, then translation */
static inline A4 FN(m4_rigid)(V3 t, const REAL q[4]) {
  V3 c0 = FN(qrot)(q, FN(v3)(R(1), R(0), R(0))), c1 = FN(qrot)(q, FN(v3)(R(0), R(1), R(0)));
  V3 c2 = FN(qrot)(q, FN(v3)(R(0), R(0), R(1)));
  A4 r = {{c0.x, c1.x, c2.x, t.x, c0.y, c1.y, c2.y, t.y, c0.z, c1.z, c2.z, t.z, R(0), R(0), R(0), R(1)}};
  return r;
}
/* getTransformationMatrix (modelDO.py:396-416) */
static inline A4 FN(m4_transformation)(REAL sx, REAL sy, REAL sz, REAL ax, REAL ay, REAL az) {
  V3 z = FN(v3)(R(0), R(0), R(0));
  REAL q0[4], qx[4], qy[4], qz[4];
  FN(quat)(z, R(0), q0);
  FN(quat)(FN(v3)(R(1), R(0), R(0)), ax, qx);
  FN(quat)(FN(v3)(R(0), R(1), R(0)), ay, qy);
  FN(quat)(FN(v3)(R(0), R(0), R(1)), az, qz);
  A4 S = FN(m4_rigid)(FN(v3)(sx, sy, sz), q0), X = FN(m4_rigid)(z, qx), Y = FN(m4_rigid)(z, qy);
  A4 Z = FN(m4_rigid)(z, qz);
  return FN(m4_mul)(S, FN(m4_mul)(X, FN(m4_mul)(Y, Z)));
}
/* applyTransformationToChild / ...Children (modelDO.py:364-394) */
static void FN(carry_child)(REAL* xf, A4 m, int parent, int child) {
  A4 P = FN(m4_load)(xf, parent), C = FN(m4_load)(xf, child);
  FN(m4_store)(xf, child, FN(m4_mul)(P, FN(m4_mul)(m, FN(m4_mul)(FN(m4_inv)(P), C))));
}
static void FN(carry_children)(REAL* xf, A4 m, int parent) {
  int env = parent / 7, kind = parent % 7;
  if (kind == 4) FN(carry_child)(xf, m, parent, env * 7 + 1);
  if (kind == 5) FN(carry_child)(xf, m, parent, env * 7 + 2);
  if (kind == 3)
    for (int i = 0; i < 7; ++i)
      if (i != kind) FN(carry_child)(xf, m, parent, env * 7 + i);
  if (kind == 0)
    for (int i = 1; i < 7; ++i)
      if (i != 3) FN(carry_child)(xf, m, parent, env * 7 + i);
}
/* transformFunc (modelDO.py:291-299) */
static void FN(transform_func)(REAL* xf, int id, A4 X, A4 m) {
  FN(carry_children)(xf, m, id);
  FN(m4_store)(xf, id, FN(m4_mul)(X, m));
}
/* rotateClampsFunc (modelDO.py:301-362) */
static void FN(rotate_clamps)(REAL* xf, int* dch, int* dchu, REAL action, int env) {
  const REAL PI = R(3.14159265358979323846);
  int rod = env * 7, left = env * 7 + 4, right = env * 7 + 5, drag = env * 3;
  REAL diff = action * R(180.0) / PI;
  if (diff == R(0)) return;
  A4 L = FN(m4_load)(xf, left), Rt = FN(m4_load)(xf, right);
  A4 local = FN(m4_mul)(FN(m4_inv)(FN(m4_load)(xf, rod)), Rt);
  REAL angle = (REAL)ASIN(local.a[4 * 2 + 1]) * R(180.0) / PI;
  int apply = 0;
  if (dch[drag] == 1 && dchu[drag] == 1) dchu[drag] = 0;
  if (action > R(0)) {
    if (angle < R(85)) {
      if (angle + diff > R(85)) diff = R(85) - angle;
      apply = 1;
    }
    if (angle > R(43) && dch[drag] == 1) { dch[drag] = 0; dchu[drag] = 1; }
  } else {
    if (angle > R(40)) {
      if (angle + diff < R(40)) diff = R(40) - angle;
      apply = 1;
    }
    if (angle < R(43) && dch[drag] == 0) { dch[drag] = 1; dchu[drag] = 1; }
  }
  if (!apply) return;
  REAL rad = diff * PI / R(180.0);
  A4 m = FN(m4_transformation)(R(0), R(0), R(0), rad, R(0), R(0));
  A4 mi = FN(m4_transformation)(R(0), R(0), R(0), -rad, R(0), R(0));
  FN(m4_store)(xf, left, FN(m4_mul)(L, mi));
  FN(carry_children)(xf, mi, left);
  FN(m4_store)(xf, right, FN(m4_mul)(Rt, m));
  FN(carry_children)(xf, m, right);
}
/* updateLaparoscopeKernel (modelDO.py:218-245) */
void FN(oracle_instrument_update)(int nEnv, const REAL* xf, const REAL* height, REAL* tips, REAL* bases) {
  for (int tid = 0; tid < 4 * nEnv; ++tid) {
    int env = tid / 4, kind = tid % 4;
    A4 X;
    V3 lo, hi;
    if (kind < 3) {
      X = FN(m4_load)(xf, env * 7 + kind);
      REAL h = height[env * 4 + kind];
      lo = FN(v3)(R(0), R(0), R(0) - h / R(1.85));
      hi = FN(v3)(R(0), R(0), R(0) + h / R(1.85));
    } else {
      X = FN(m4_load)(xf, env * 7);
      REAL rod = height[env * 4], clamp = height[env * 4 + 2];
      lo = FN(v3)(R(0), R(0), R(0) - rod * R(0.5) - clamp * R(1.0));
      hi = FN(v3)(R(0), R(0), R(0) + rod * R(0.5) + clamp * R(1.0));
    }
    FN(st)(bases, tid, FN(m4_point)(X, lo));
    FN(st)(tips, tid, FN(m4_point)(X, hi));
  }
}
/* applyActionsKernel (modelDO.py:175-199), threads in tid order, then
 * updateLaparoscopeKernel; actions [nEnv*5] */
void FN(oracle_instrument_actions)(int nEnv, REAL* xf, const REAL* height, int* dch, int* dchu,
                                   const REAL* actions, REAL* tips, REAL* bases) {
  for (int tid = 0; tid < 7 * nEnv; ++tid) {
    int env = tid / 7, kind = tid % 7;
    A4 X = FN(m4_load)(xf, tid);
    const REAL* a = actions + 5 * env;
    if (kind == 0)
      FN(transform_func)(xf, tid, X, FN(m4_transformation)(R(0), R(0), a[0], R(0), R(0), a[1]));
    else if (kind == 3)
      FN(transform_func)(xf, tid, X, FN(m4_transformation)(R(0), R(0), R(0), a[2], a[3], R(0)));
    else if (kind == 4)
      FN(rotate_clamps)(xf, dch, dchu, a[4], env);
  }
  FN(oracle_instrument_update)(nEnv, xf, height, tips, bases);
}
/* transfromCartesianActionsKernel[InWorkspace] (modelDO.py:55-154), then
 * updateLaparoscopeKernel; lo/hi NULL for the unbounded form */
void FN(oracle_instrument_cartesian)(int nEnv, REAL* xf, const REAL* height, const REAL* actions,
                                     const REAL* lo, const REAL* hi, int* violation, REAL* tips, REAL* bases) {
  for (int env = 0; env < nEnv; ++env) {
    REAL av[3] = {actions[3 * env], actions[3 * env + 1], actions[3 * env + 2]};
    if (av[0] == R(0) && av[1] == R(0) && av[2] == R(0)) continue;
    V3 cur = FN(ld)(tips, 4 * env + 3);
    if (lo && hi) {
      violation[env] = 0;
      REAL nw[3] = {cur.x + av[0], cur.y + av[1], cur.z + av[2]};
      for (int c = 0; c < 3; ++c)
        if (nw[c] < lo[c] || nw[c] > hi[c]) { av[c] = R(0); violation[env] = 1; }
    }
    A4 rodS = FN(m4_T)(FN(m4_load)(xf, 7 * env));      /* stored (transposed) forms, as the kernel reads them */
    A4 rcS = FN(m4_T)(FN(m4_load)(xf, 7 * env + 3));
    V3 rel = FN(m4_point)(FN(m4_T)(FN(m4_inv)(rcS)), cur);
    V3 delta = FN(m4_point)(rcS, FN(v3)(av[0], av[1], av[2]));
    V3 nrel = FN(add)(rel, delta);
    REAL r0 = FN(length)(rel), r1 = FN(length)(nrel);
    V3 vc = FN(cross)(rel, nrel);
    REAL sinT = FN(length)(vc) / (FN(length)(rel) * FN(length)(nrel));
    REAL q[4];
    FN(quat)(FN(normalize)(vc), (REAL)ASIN(sinT), q);
    FN(transform_func)(xf, 7 * env, FN(m4_T)(rodS), FN(m4_transformation)(R(0), R(0), r1 - r0, R(0), R(0), R(0)));
    FN(transform_func)(xf, 7 * env + 3, FN(m4_T)(rcS), FN(m4_rigid)(FN(v3)(R(0), R(0), R(0)), q));
  }
  FN(oracle_instrument_update)(nEnv, xf, height, tips, bases);
}
/* setRotationCenterTransformKernel (modelDO.py:156-173), then updateLaparoscopeKernel */
void FN(oracle_instrument_set_rcm)(int nEnv, REAL* xf, const REAL* height, const REAL* t, const REAL* r,
                                   REAL* tips, REAL* bases) {
  for (int env = 0; env < nEnv; ++env) {
    A4 X = FN(m4_load)(xf, 7 * env + 3);
    A4 m = FN(m4_transformation)(t[3 * env], t[3 * env + 1], t[3 * env + 2], r[3 * env], r[3 * env + 1], r[3 * env + 2]);
    FN(transform_func)(xf, 7 * env + 3, X, FN(m4_mul)(FN(m4_inv)(X), m));
  }
  FN(oracle_instrument_update)(nEnv, xf, height, tips, bases);
}
/* 4x4 helpers exported for the tests */
void FN(oracle_m4_inverse)(const REAL* m, REAL* out) {
  A4 x;
  for (int i = 0; i < 16; ++i) x.a[i] = m[i];
  A4 r = FN(m4_inv)(x);
  for (int i = 0; i < 16; ++i) out[i] = r.a[i];
}
void FN(oracle_m4_transformation)(REAL sx, REAL sy, REAL sz, REAL ax, REAL ay, REAL az, REAL* out) {
  A4 r = FN(m4_transformation)(sx, sy, sz, ax, ay, az);
  for (int i = 0; i < 16; ++i) out[i] = r.a[i];
}

#undef A4
#undef V3
#undef FEPS
#undef R
#undef FN
