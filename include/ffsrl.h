/*
 * ffsrl.h - C ABI of the MI355X (gfx950) implementation of FF-SRL's per-step
 * simulation path: XPBD/PBD predict and integrate, distance / volume /
 * Neo-Hookean tetrahedral constraint projection with the grasp constraint,
 * triangle-BVH refit and the per-environment ray-cast renderer.
 *
 * Conventions
 *   - Every entry point returns 0 on success, a hipError_t value (> 0) for a
 *     HIP failure, or a negative FFSRL_E_* code; ffsrl_last_error() gives the
 *     message (thread-local).  No C++ exception crosses this boundary.
 *   - Device pointers are non-owning; the caller (the Python layer, through
 *     torch) owns every buffer.  `stream` is a hipStream_t passed as void*;
 *     all work is stream-ordered on it and nothing synchronises the host,
 *     so every entry point except ffsrl_bvh_build may be captured into a HIP
 *     graph.
 *   - Positions are float4 (x, y, z, inverse mass); velocities float4 (vx, vy,
 *     vz, 0).  Topology uses mesh-local vertex ids; an environment adds its
 *     vertex base.  Environments that share a mesh share its topology arrays.
 */
#ifndef FFSRL_H
#define FFSRL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFSRL_ABI_VERSION 14

#define FFSRL_E_ARG (-1)      /* invalid argument */
#define FFSRL_E_LIMIT (-2)    /* a size exceeds a documented kernel limit */
#define FFSRL_E_STATE (-3)    /* call made in the wrong state */

/* Per-env status bits of the batched and shared frames (ffsrl_batch.status),
 * sticky until the caller clears them. */
#define FFSRL_STATUS_SATURATED 1u   /* a Jacobi term beyond the fixed-point range, or not finite:
                                       clamped, so the frame no longer equals the plain form */
#define FFSRL_STATUS_INCOMPLETE 2u  /* a shared frame could not complete (a bounded wait gave up) */

/* Most Jacobi terms one vertex may receive in one pass of the batched frame
 * (ffsrl_env_desc.max_terms): each term is clamped to 2^24 - 1 quanta, so
 * 128 of them cannot overflow the 32-bit fixed-point sums. */
#define FFSRL_MAX_TERMS 128

/* Constraint modes (per constraint family). */
#define FFSRL_OFF 0
#define FFSRL_PBD 1   /* stiffness form: distanceConstraints / volumeConstraints */
#define FFSRL_XPBD 2  /* compliance form: *ConstraintsXPBD */

/* One environment of the CSR environment table (80 bytes, int32 fields).
 * Ranges index the shared topology arrays of ffsrl_batch; state ranges index
 * the per-environment state arrays. */
typedef struct ffsrl_env_desc {
  int32_t vbase, nv;       /* state vertices [vbase, vbase+nv) */
  int32_t ebase, ne;       /* edges of the env's mesh */
  int32_t tbase, nt;       /* tets of the env's mesh */
  int32_t fbase, nf;       /* render triangles of the mesh, BVH leaf order */
  int32_t nodebase, nn;    /* BVH topology nodes of the mesh */
  int32_t boundbase;       /* first BVH bound (state) of this env */
  int32_t lvlbase, nlvl;   /* refit groups level_off[lvlbase .. lvlbase+nlvl]: all leaves, then
                              inner nodes level by level, deepest first */
  int32_t lam_ebase;       /* first edge lambda of this env (XPBD, iterations > 1) */
  int32_t lam_tbase;       /* first tet lambda of this env */
  int32_t dp_exp;          /* quantum of the fixed-point Jacobi sums of the batched
                              frame: 2^dp_exp (the host mirror uses 2^-26 x bbox diagonal,
                              rounded down to a power of two: a term's range is D/8 .. D/4) */
  int32_t vtopo_base;      /* first vf_off entry of the env's mesh */
  int32_t bvh_depth;       /* levels of the env's BVH (max node level + 1) */
  int32_t max_terms;       /* most Jacobi terms of one vertex in one pass: incident tets plus
                              incident edges, or incident surface triangles with collision;
                              the batched frames reject envs above FFSRL_MAX_TERMS */
  int32_t reserved2;
} ffsrl_env_desc;

/* Device arrays of a batch of environments. */
typedef struct ffsrl_batch {
  int32_t n_env;
  int32_t max_nv;                 /* largest nv over envs (sizes LDS) */
  int64_t n_tet_total;            /* length of the tet arrays (plane stride of tet_inv_rest) */
  const ffsrl_env_desc* envs;     /* [n_env] */
  /* topology (shared per mesh) */
  const int32_t* edges;           /* int2 [E]: mesh-local (a, b) */
  const float* edge_rest;         /* [E] */
  const int32_t* tets;            /* int4 [T]: mesh-local (a, b, c, d) */
  const uint64_t* tets16;         /* [T]: the same ids packed as 4 x 16 bits (a in the low bits);
                                     read by the batched frame */
  const uint32_t* edges16;        /* [E]: edge ids packed as 2 x 16 bits; read by the batched frame */
  const float* tet_rest_volume;   /* [T] */
  const float* tet_inv_rest;      /* [9][T] planes of the row-major inverse rest shape matrix */
  const int32_t* tris;            /* int4 [F]: mesh-local (a, b, c, original face id), leaf order */
  const int32_t* node_left;       /* [N] left child (mesh-local node id) or -1 for a leaf */
  const int32_t* node_first;      /* [N] first leaf-order triangle of a leaf */
  const int32_t* node_count;      /* [N] triangle count of a leaf, 0 for inner nodes */
  const int32_t* level_nodes;     /* [N] node ids in refit-group order */
  const int32_t* level_info;      /* int2 [N] per level_nodes entry: (left child or first triangle,
                                     triangle count) */
  const int32_t* level_off;       /* per env lvlbase: offsets into level_nodes (relative to nodebase) */
  const int32_t* vf_off;          /* per mesh [nv+1] at vtopo_base: CSR offsets of each vertex's faces */
  const int32_t* vf_idx;          /* per mesh [3F] at 3*fbase: leaf positions of a vertex's faces,
                                     in increasing original face id */
  /* textures of the render mesh (colour mode 1); NULL when no mesh has one */
  const float* tri_uv;            /* [F][6] per render triangle (leaf order): texture coordinates
                                     (s, t) of its corners a, b, c */
  const int32_t* tri_tex;         /* [F] texture of each render triangle (leaf order), or -1: the
                                     mean of its vertex colours */
  const float* texels;            /* float4 [n_texels]: RGB texels of every texture, the texels of
                                     each square texture consecutive */
  const int64_t* tex_info;        /* [n_tex][2]: first texel and side of each texture */
  int64_t n_texels;
  /* state (per environment) */
  float* x;                       /* float4 [V] positions + inverse mass */
  float* p;                       /* float4 [V] predicted positions + inverse mass */
  float* v;                       /* float4 [V] velocities */
  float* drag;                    /* [V] grasp weight (activeDragConstraint) */
  float* lam_e;                   /* [sum E] XPBD edge lambdas (NULL when not carried) */
  float* lam_t;                   /* [sum T] XPBD tet lambdas (NULL when not carried) */
  float* bvh_lo;                  /* float4 [sum N]: aabb min, .w = int bits (first tri or left child) */
  float* bvh_hi;                  /* float4 [sum N]: aabb max, .w = int bits (triangle count) */
  float* tips;                    /* [n_env*4*3] instrument points; grasp pulls to point 2 */
  float* caps_base;               /* [n_env*4*3] instrument capsule bases (tips are `tips`), or NULL */
  const float* caps_radius;       /* [n_env*4] instrument capsule radii */
  int32_t* in_collision;          /* [n_env] surface-triangle contacts, summed over passes until cleared */
  uint32_t* sched;                /* [2] work counters of ffsrl_frame (NULL: one workgroup per env);
                                     zero before the first frame, left zero by every frame */
  /* plain-form scratch (per-kernel launches with global atomics) */
  float* dp;                      /* [3][V] planes (NULL unless the plain form is used) */
  int32_t* cnt;                   /* [V] */
  /* shared-env frame scratch (ffsrl_frame_shared; NULL otherwise) */
  uint64_t* split_acc;            /* [V][2] fixed-point Jacobi sums; zero before the first
                                     frame, left zero by every frame */
  uint32_t* split_ctr;            /* [n_env][4] chunk counters (taken, done, give-up flag, exits);
                                     zero before the first frame, left zero by every frame */
  int32_t* status;                /* [n_env] FFSRL_STATUS_* bits, or'ed by the batched and shared
                                     frames (required by them); the caller clears them */
  uint64_t* fold_ws;              /* workspace of the shared frame's folded form: 7 planes of
                                     fold_stride x 16 B (3 sum planes, zero before the first frame
                                     and left at zero by every frame; 2 + 2 state planes); envs of
                                     at most 5,104 vertices without collision take the folded form
                                     when it is given (NULL: none does) */
  int64_t fold_stride;            /* vertices per plane of fold_ws (>= every env's vbase + nv) */
  /* host mirrors needed by the plain form, which launches per env (and per
   * BVH level) from the host: the env table and the level offsets */
  const ffsrl_env_desc* host_envs;   /* [n_env] (host memory) */
  const int32_t* host_level_off;     /* same content as level_off (host memory) */
} ffsrl_batch;

/* Scalars of SimModelDO handed to the kernels (modelDO.py:774-830). */
typedef struct ffsrl_params {
  int32_t substeps, iterations;
  float dt;                       /* substep dt = (1/fps)/substeps */
  float gravity[3];
  float damping;                  /* velocityDampening */
  float ground;                   /* groundLevel */
  int32_t dist_mode;              /* FFSRL_OFF / PBD / XPBD */
  float dist_k;                   /* kS (PBD) or alpha~ = compliance/dt^2 (XPBD) */
  int32_t vol_mode;
  float vol_k;
  int32_t nh_on;                  /* Neo-Hookean on/off */
  float nh_vol_c, nh_dev_c;       /* hydrostatic / deviatoric compliance (1/lambda, 1/mu) */
  float drag_ks;                  /* globalKsDrag */
  int32_t lambda_legacy;          /* 1: zero lambdas only when iterations > 1 */
  int32_t collide;                /* 1: instrument collision after the grasp (collideLaparoscopeDO) */
  float contact_k;                /* globalKsContact */
} ffsrl_params;

/* Camera of the ray-cast renderer: one origin and a precomputed direction per
 * ray of the (2H x 2W) supersampled ray grid (render.py:299-311). */
typedef struct ffsrl_camera {
  float origin[3];
  const float* dirs;              /* float4 [2H*2W] (device), row-major ray order */
  int32_t height, width;          /* output pixels per env; width % 4 == 0, height % 4 == 0 */
  int32_t mode;                   /* 4 depth, 3 vertex colour, 2 normal, 0 grayscale, 1 texture (vertex
                                     colours where a triangle has none), 5 per-triangle colour */
  float light_dir[3];
  float light_intensity;
  int32_t traversal;              /* FFSRL_TRAVERSE_*: 0 (zero-initialised) picks by image size */
} ffsrl_camera;

/* Ray traversal of the renderer.  Both forms give the same hits, depths and
 * colours bit for bit; they differ only in speed.  PER_RAY: a lane traces
 * one supersampled ray.  QUAD: a lane traces the 2x2 rays of its pixel with
 * one stack (a BVH of at most 2^28 - 1 nodes).  AUTO: QUAD from 128x128
 * output pixels, PER_RAY below. */
#define FFSRL_TRAVERSE_AUTO 0
#define FFSRL_TRAVERSE_PER_RAY 1
#define FFSRL_TRAVERSE_QUAD 2

/* One tick of serving requests (device arrays of n entries, env ids distinct). */
#define FFSRL_REQ_STEP 0
#define FFSRL_REQ_RESET 1
typedef struct ffsrl_requests {
  int32_t n;
  const int32_t* envs;            /* env id of each request */
  const int32_t* kind;            /* FFSRL_REQ_STEP / FFSRL_REQ_RESET */
  const float* action;            /* float4 per request: tool delta xyz, clamp (-1 keep, 0 open, 1 close) */
  const float* x0;                /* float4 [V]: rest positions (xyz) and inverse masses (w) */
  const float* tips0;             /* float [n_env*4*3]: initial instrument points */
  const float* lo;                /* float[3] workspace lower corner, or NULL */
  const float* hi;                /* float[3] workspace upper corner, or NULL */
  const float* base0;             /* float [n_env*4*3]: initial capsule bases (reset), or NULL */
} ffsrl_requests;

int ffsrl_abi_version(void);
const char* ffsrl_last_error(void);

/* ---- batched form ------------------------------------------------------ */

/* One frame of every environment in one launch: S substeps x I iterations of
 * predict, projection (volume, distance, Neo-Hookean, Jacobi-averaged), grasp,
 * ground clamp and integrate, followed by the BVH refit when refit != 0.
 * Replaces SimIntegratorDO.stepModel (integrator.py:1640-1744) followed by
 * SimBVH.refitBVH (bvh.py:442-451).  `order` (device, may be NULL) gives the
 * dispatch order of environments (largest first balances mixed sizes).
 * block = threads per environment (256, 512 or 1024).  The Jacobi sums are
 * fixed point: an env whose terms leave the range (or are not finite) gets
 * FFSRL_STATUS_SATURATED in status[env].  Envs above FFSRL_MAX_TERMS (read
 * from host_envs when given) are rejected with FFSRL_E_LIMIT. */
int ffsrl_frame(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* order, int32_t block,
                int32_t refit, void* stream);

/* ffsrl_frame restricted to the n environments listed in envs (device, ids
 * distinct, in dispatch order): the serving loop steps only the scenes with a
 * pending step request.  Same reference counterpart as ffsrl_frame. */
int ffsrl_frame_envs(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* envs, int32_t n, int32_t block,
                     int32_t refit, void* stream);

/* ffsrl_frame_envs with every CU working on the n listed environments:
 * each environment's frame is cut into chunks (constraint ranges of a
 * projection pass, vertex ranges of the apply passes, triangle ranges of the
 * collision pass) that the workgroups on one XCD take from the
 * environment's counter, passes ordered by its completion counter.  For
 * batches with fewer environments than CUs (serving ticks, few large
 * scenes).  Results equal ffsrl_frame_envs bit for bit (integer Jacobi sums).
 * envs NULL: environments 0 .. n-1.
 * A workgroup only waits for chunks that running workgroups hold, so the
 * launch needs no co-residency; waits are bounded, and a frame that could
 * not complete sets FFSRL_STATUS_INCOMPLETE in status[env] (its global
 * sums are cleared and its counters reset, so the next frame starts
 * clean).  Launches that overlap on one environment race.  Same reference
 * counterpart as ffsrl_frame. */
int ffsrl_frame_shared(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* envs, int32_t n,
                       int32_t block, int32_t refit, void* stream);

/* ffsrl_frame_shared for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * global sums and chunk counters live in the library's own device
 * workspace, 32 MiB of sums (2,097,152 vertex slots of 16 B) plus counters
 * for 4096 listed envs, one copy per device held until the process ends,
 * instead of b->split_acc / b->split_ctr (ignored).  Listed env el uses
 * vertex slots [el * max_nv, (el + 1) * max_nv): n * max_nv must stay
 * within 2,097,152 and n within 4096, else FFSRL_E_LIMIT.  Nothing is
 * allocated: call ffsrl_xcd_count() once per device beforehand (it
 * allocates; FFSRL_E_STATE otherwise).  Every frame leaves the workspace at
 * zero; calls that overlap on one device race.  Results equal
 * ffsrl_frame_shared bit for bit.  No reference counterpart (the reference
 * allocates through Warp). */
int ffsrl_frame_shared_noalloc(const ffsrl_batch* b, const ffsrl_params* prm, const int32_t* envs, int32_t n,
                               int32_t block, int32_t refit, void* stream);

/* Partition of each environment's constraints for ffsrl_frame_parts (built
 * on the host from the current constraint order; rebuild it when that order
 * changes).  A part is a range of its mesh's tets and a range of its edges;
 * its vertices are numbered locally. */
#define FFSRL_PART_OWNER 0x80000000u   /* verts entry: this part writes the vertex back */
#define FFSRL_PART_SHARED 0x40000000u  /* verts entry: other parts hold the vertex too */
typedef struct ffsrl_parts {
  int32_t n_parts;
  int32_t max_local;              /* most local vertices of a part of an env with several parts
                                     (sizes LDS); at most ffsrl_parts_max_local() */
  int32_t max_solo;               /* most vertices of an env held by one part (its vertex list is the
                                     identity; sizes LDS); at most ffsrl_parts_max_solo() */
  int32_t reserved;
  const int32_t* parts;           /* int [n_parts][8]: tet begin, tet end, edge begin, edge end (within
                                     the mesh's tets / edges, current order), first verts entry, local
                                     vertex count, first tris16 entry, triangle count (collision) */
  const int32_t* env_parts;       /* int2 [n_env]: first part and part count of each env (its mesh's) */
  const int32_t* verts;           /* per part, its local vertices in local-id order: mesh-local id,
                                     | FFSRL_PART_SHARED, | FFSRL_PART_OWNER on exactly one part */
  const uint64_t* tets16;         /* [T] each tet's ids local to its part, 4 x 16 bits (a low) */
  const uint32_t* edges16;        /* [E] each edge's ids local to its part, 2 x 16 bits */
  const uint64_t* tris16;         /* by part, its surface triangles' local ids, 3 x 16 bits (collision
                                     pass; NULL without it) */
  uint64_t* sums;                 /* [3][stride][2] fixed-point sums of shared vertices, by state
                                     vertex; zero before the first frame, left zero by every frame */
  int64_t stride;                 /* vertices per plane of sums (>= every env's vbase + nv) */
  uint32_t* ctr;                  /* [4 n_env + 64]: counters in the first 4 n_env + 17 words (zero
                                     before the first frame, left zero by every frame), then the
                                     claim table's extent per XCD (written by every launch) */
  int32_t* claims;                /* [n_claims] workspace: each launch's claim table */
  int64_t n_claims;               /* >= the total part count of all envs */
  const int32_t* host_env_parts;  /* host mirror of env_parts (launch checks), required */
} ffsrl_parts;

/* ffsrl_frame_envs with each listed environment cut into parts, one
 * workgroup (1024 threads) per part for the whole frame: positions and sums
 * of a part's vertices stay in LDS, and only vertices shared between parts
 * exchange sums through global memory, once per pass (the projection and,
 * with prm->collide, the collision pass over each part's surface triangles,
 * parts->tris16).  Results equal ffsrl_frame_envs bit for bit (integer
 * Jacobi sums), contact counts included; envs larger than the batched
 * frame's ffsrl_frame_max_vertices() run too, when their parts fit
 * (ffsrl_parts_max_local()).  All parts of an env run on one XCD at once:
 * at most CUs / XCDs (and 63) parts per env (FFSRL_E_LIMIT).  Two
 * launches: the claim table, then the frame.  The launch should be the
 * only persistent frame on the device: waits are bounded, and a frame that
 * could not complete sets FFSRL_STATUS_INCOMPLETE (its sums are cleared and
 * its counters reset).  envs NULL: environments 0 .. n-1.  Same reference
 * counterpart as ffsrl_frame. */
int ffsrl_frame_parts(const ffsrl_batch* b, const ffsrl_params* prm, const ffsrl_parts* parts,
                      const int32_t* envs, int32_t n, int32_t refit, void* stream);

/* Most local vertices of one part of an env cut into several parts (56 B of
 * LDS each), and most vertices of an env held by one part (32 B each). */
int ffsrl_parts_max_local(void);
int ffsrl_parts_max_solo(void);

/* Number of XCDs of the current device (probed once with one workgroup per
 * CU; the shared frame assigns environments to XCDs).  Allocates on its
 * first call per device, so call it before capturing a shared frame into a
 * graph.  No reference counterpart (launch-shape query). */
int ffsrl_xcd_count(void);

/* Largest nv that ffsrl_frame can hold in LDS (dP + counters). */
int ffsrl_frame_max_vertices(void);

/* Bottom-up refit of every environment's BVH.  Replaces refitBVHLoop
 * (bvh.py:419-440) / refitBVHLevel (bvh.py:30-104). */
int ffsrl_bvh_refit(const ffsrl_batch* b, int32_t env_begin, int32_t env_end, void* stream);

/* Ray-cast render of environments [env_begin, env_end) with the 2x2 box
 * filter.  Colour mode 1 is the reference's texture mode (render.py:356-380):
 * the hit's texture coordinates u s_a + v s_b + w s_c (batch tri_uv) looked
 * up bilinearly in its triangle's texture (texture_interpolation,
 * render.py:124-148, with its y + x * width indexing; reads past the texel
 * table give zero), or the mean of the triangle's vertex colours (vcolor,
 * required) where it has no texture; mode 5 uses per-triangle colours
 * (tcolor, leaf order).
 * depth: [n][H][W] (mode 4) ; rgb: [n][H][W][3] (other modes);
 * a colour mode given a depth output too writes both from one traversal
 * (RGB-D), each equal to what its own mode writes;
 * hit / thit (optional, may be NULL): per-ray face ids and t, [n][2H][2W].
 * Replaces drawSimBVHKernelManyEnvs + meshQueryRayManyEnvs + downsample_kernel
 * (render.py:276-498, 699-724) launched by renderFunctionNew (render.py:975-1022).
 * Output env index = env - env_begin. */
int ffsrl_render(const ffsrl_batch* b, const ffsrl_camera* cam, int32_t env_begin, int32_t env_end,
                 float* depth, float* rgb, const float* normals, const float* vcolor,
                 const float* tcolor, int32_t* hit, float* thit, void* stream);

/* ffsrl_render of the n environments listed in envs (device, ids distinct).
 * Outputs are indexed by env id (they hold n_env images), so each scene of
 * the serving loop keeps its latest observation in place.  Same reference
 * counterpart as ffsrl_render. */
int ffsrl_render_envs(const ffsrl_batch* b, const ffsrl_camera* cam, const int32_t* envs, int32_t n,
                      float* depth, float* rgb, const float* normals, const float* vcolor,
                      const float* tcolor, int32_t* hit, float* thit, void* stream);

/* Area-weighted vertex normals (float4 [V]) of the render mesh.  Replaces
 * zeroVec3Array + vertex_normal_sum_kernel + normalize_kernel (render.py:81-121).
 * Each vertex sums the area normals of its faces in increasing face id (the
 * vf_off / vf_idx adjacency), the order in which the reference's atomics
 * accumulate when its threads run in tid order, so the sums are deterministic. */
int ffsrl_vertex_normals(const ffsrl_batch* b, int32_t env_begin, int32_t env_end, float* normals,
                         void* stream);

/* Per-env tool update: adds delta[env] (3 floats) to the 4 instrument points
 * of each env, less the components that would take the drag point (point 3)
 * out of [lo, hi] (lo, hi device float[3] each, or NULL), which are dropped
 * as transfromCartesianActionsKernelInWorkspace drops them
 * (modelDO.py:97-154); capsule bases, when present, move with their tips.
 * Stands in for applyCartesianActionsInWorkspace (modelDO.py:1385-1404) with
 * the tool translated directly (no remote-centre-of-motion kinematics). */
int ffsrl_move_tools(const ffsrl_batch* b, const float* delta, const float* lo, const float* hi,
                     void* stream);

/* Grasp selection: drag[v] = 1 if |x[v] - tip[env*4+3]| < radius for envs with
 * state[env*3] == 1 and update[env*3] != 0, else 0 (updateClampsDragKernel,
 * modelDO.py:247-279).  state/update are int32 [n_env*3] device arrays. */
int ffsrl_grasp_select(const ffsrl_batch* b, float radius, const int32_t* state,
                       const int32_t* update, void* stream);

/* ---- instrument kinematics ---------------------------------------------- */
/* The instrument of each env (laparoscope.py:375-386): seven 4x4 transforms
 * per env, stored transposed (row-vector convention, as the reference keeps
 * them), at xform[(env*7 + k)*16]:
 *   0 rod capsule, 1 left jaw, 2 right jaw, 3 rotation centre (the remote
 *   centre of motion), 4 left jaw pivot, 5 right jaw pivot, 6 rod.
 * A transform applied to 3 carries every other one, applied to 0 carries
 * 1, 2, 4, 5, 6, applied to 4 carries 1 and to 5 carries 2
 * (applyTransformationToChildren, modelDO.py:364-394).  height holds the
 * capsule heights (rod, left jaw, right jaw, 0) per env; drag_cut_heat and
 * drag_cut_heat_update the clamp state (3 ints per env, drag first);
 * violation the workspace flags (NULL when unused).  The capsule points
 * derived from the transforms are written to the batch's tips and
 * caps_base. */
typedef struct ffsrl_instrument {
  int32_t n_env;
  float* xform;                   /* [n_env*7*16] */
  const float* height;            /* [n_env*4] */
  int32_t* drag_cut_heat;         /* [n_env*3] */
  int32_t* drag_cut_heat_update;  /* [n_env*3] */
  int32_t* violation;             /* [n_env] or NULL */
} ffsrl_instrument;

/* applyActions (modelDO.py:1331-1362): actions float[n_env*5] per env (rod
 * insertion, rod roll, rotation-centre pitch, yaw, jaw opening in radians);
 * applyActionsKernel (:175-199) with its three transform threads of an env
 * run in thread order (rod, rotation centre, jaws), the jaw state machine
 * rotateClampsFunc (:301-362), updateLaparoscopeKernel (:218-245) into
 * tips / caps_base, then updateClampsDragKernel (:247-279) with
 * lookup_radius (lookup_radius < 0 skips it, as the jaw animation of
 * forceLaparoscopeClampVertex does, :1533-1568). */
int ffsrl_instrument_actions(const ffsrl_batch* b, const ffsrl_instrument* ins, const float* actions,
                             float lookup_radius, void* stream);
/* applyCartesianActions / applyCartesianActionsInWorkspace (modelDO.py:
 * 1364-1404; transfromCartesianActionsKernel[InWorkspace], :55-154): moves
 * the drag point (tip[env*4+3]) by actions[env*3..+2] about the rotation
 * centre, then updates tips / caps_base.  With lo and hi (device float[3]
 * each) the components that would take the drag point out of the box are
 * dropped and violation[env] is set to 1 (0 otherwise) for envs with a
 * nonzero action. */
int ffsrl_instrument_cartesian(const ffsrl_batch* b, const ffsrl_instrument* ins, const float* actions,
                               const float* lo, const float* hi, void* stream);
/* setRotationCenter (modelDO.py:1570-1587; setRotationCenterTransformKernel,
 * :156-173): the rotation centre of env e becomes
 * T(translation[e]) Rx Ry Rz(rotation[e]) (radians), carrying the rest of the
 * instrument; then updates tips / caps_base. */
int ffsrl_instrument_set_rcm(const ffsrl_batch* b, const ffsrl_instrument* ins, const float* translation,
                             const float* rotation, void* stream);
/* updateLaparoscopeKernel (modelDO.py:218-245) alone. */
int ffsrl_instrument_update(const ffsrl_batch* b, const ffsrl_instrument* ins, void* stream);

/* ---- serving ------------------------------------------------------------ */

/* Applies one tick of requests: a reset restores the env like resetFixedModel
 * (modelDO.py:1589-1622: positions from x0, velocities and grasp weights 0,
 * instrument points from tips0), refits its BVH and clears the env's status
 * word (when the batch has one: a reset starts a new episode); a step adds the action's
 * delta to the env's 4 instrument points, less the components that would
 * take the drag point out of [lo, hi] when given (dropped, as
 * transfromCartesianActionsKernelInWorkspace does, modelDO.py:97-154), and, when
 * the clamp command is 0 or 1, reselects the grasp like updateClampsDragKernel
 * (modelDO.py:247-279): drag = |x - tip[env*4+3]| < radius when closed, 0 when
 * open. */
int ffsrl_apply_requests(const ffsrl_batch* b, const ffsrl_requests* r, float radius, void* stream);
/* (With capsule bases in the batch, a step moves each base by its tip's
 * clamped displacement and a reset restores base0 when given.) */

/* 32 observation scalars per listed env, written to out[env*32 + k]:
 *   0-11 the 4 instrument points, 12-14 centroid, 15-17 / 18-20 bounding box
 *   min / max, 21-23 mean velocity, 24 max speed, 25 grasped vertices,
 *   26-28 centroid of the grasped vertices (0 if none), 29 vertices at or
 *   below the ground level, 30 / 31 max / mean displacement from x0.
 * Sums use a fixed reduction tree (deterministic).  No reference counterpart
 * (the reference's observations are read on the host from torch views,
 * modelDO.py:1483-1518). */
int ffsrl_observe_envs(const ffsrl_batch* b, const int32_t* envs, int32_t n, const float* x0, float ground,
                       float* out, void* stream);

/* ---- plain form: one launch per kernel per environment ------------------ */
/* The reference's per-kernel launch model (integrator.py:1644-1741), restricted
 * to environments [env_begin, env_end) with one launch per environment. */
int ffsrl_predict(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end, void* stream);
int ffsrl_project_volume(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end, void* stream);
int ffsrl_project_distance(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end, void* stream);
int ffsrl_project_neohookean(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end, void* stream);
int ffsrl_apply(const ffsrl_batch* b, int32_t env_begin, int32_t env_end, void* stream);
int ffsrl_grasp(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end, void* stream);
/* collideLaparoscopeDO (integrator.py:981-1050, launched at :1713-1726):
 * every surface triangle against instrument capsules 0..2 of its env, the
 * responses subtracted from p for vertices with inverse mass > 0 after all
 * triangles have read p; in_collision[env] += triangles with a response. */
int ffsrl_collide(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end, void* stream);
int ffsrl_bounds_integrate(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end, void* stream);
int ffsrl_zero_lambdas(const ffsrl_batch* b, int32_t env_begin, int32_t env_end, void* stream);
/* One launch per env and BVH level, deepest level first (refitBVHLoop). */
int ffsrl_bvh_refit_plain(const ffsrl_batch* b, int32_t env_begin, int32_t env_end, void* stream);

/* One frame in the reference's own launch model: SimIntegratorDO.stepModel
 * (integrator.py:1640-1744) launching each DO kernel once per substep and
 * iteration over every environment, then refitBVHLoop (bvh.py:419-454) as one
 * launch per refit group over every environment.  The kernels are the plain
 * form's (global float atomics; results equal the per-env plain form up to
 * the order of the atomics); envs env_begin .. env_end-1 must have
 * contiguous vertices (and lambdas when carried).  Needs dp/cnt and the host
 * mirrors; keeps device prefix sums of the envs' counts (re-uploaded when
 * the counts change), so it synchronises the host on its first call for an
 * env range and is not for graph capture then. */
int ffsrl_frame_do(const ffsrl_batch* b, const ffsrl_params* prm, int32_t env_begin, int32_t env_end,
                   int32_t refit, void* stream);

/* remapAToBKernel (utils.py:32-42): dst[map[2i]] = src[map[2i+1]] for float4
 * arrays, i < n. */
int ffsrl_remap(float* dst, const float* src, const int32_t* map, int64_t n, void* stream);

/* ---- host: BVH construction -------------------------------------------- */
/* Binned-SAH build of one mesh (SimBVH.buildBVH / subdivide /
 * findBestSplitPlaneBinned, bvh.py:281-518), host memory.  Outputs hold at
 * least 2*nf entries (tri_map: nf); the node count and deepest level are
 * returned through num_nodes and max_level. */
int ffsrl_bvh_build(const float* verts, int32_t nv, const int32_t* faces, int32_t nf, int32_t bins,
                    int32_t* left, int32_t* first, int32_t* count, int32_t* level, int32_t* tri_map,
                    int32_t* num_nodes, int32_t* max_level);

/* Device build of the same topology (ffsrl_bvh_build's outputs, node for
 * node), one workgroup per job; all pointers of a job are device memory and
 * `jobs` itself is a device array.  info[0] = node count (-1 if the tree was
 * deeper than 64 levels), info[1] = deepest level.  scratch holds
 * ffsrl_bvh_scratch_bytes(nf) bytes.  Replaces SimBVH.buildBVH
 * (bvh.py:456-518) for many meshes at once (scene admission). */
typedef struct ffsrl_bvh_job {
  const float* verts;             /* [nv*3] */
  const int32_t* faces;           /* [nf*3] */
  int32_t nv, nf;
  int32_t *left, *first, *count, *level;  /* [2*nf] each */
  int32_t* tri_map;               /* [nf] */
  int32_t* info;                  /* [2] */
  void* scratch;
} ffsrl_bvh_job;
int64_t ffsrl_bvh_scratch_bytes(int32_t nf);
int ffsrl_bvh_build_device(const ffsrl_bvh_job* jobs, int32_t n_jobs, int32_t bins, void* stream);

/* Greedy colouring of n constraints of `arity` vertices (host memory): two
 * constraints that share a vertex get different colours.  No reference
 * counterpart (the reference accumulates with global atomics in element
 * order, integrator.py:105-109). */
int ffsrl_color_constraints(const int32_t* elems, int64_t n, int32_t arity, int32_t nv, int32_t* color,
                            int32_t* num_colors);

/* Order in which the frame kernel visits a mesh's constraints (host memory):
 * sorted along a Morton curve of the constraint centroids (verts: nv x 3
 * floats), then, within each run of `chunk` consecutive constraints, stably
 * by a greedy colour so that the lanes of one LDS atomic rarely share a
 * vertex.  perm[i] = element visited i-th.  The frame's integer Jacobi sums
 * make the result independent of this order.  No reference counterpart. */
int ffsrl_order_constraints(const int32_t* elems, int64_t n, int32_t arity, const float* verts, int32_t nv,
                            int32_t chunk, int64_t* perm);

/* Partition of one mesh for ffsrl_frame_parts (host memory): part p takes
 * tets [p nt / parts, (p+1) nt / parts) and edges [p ne / parts, (p+1) ne /
 * parts) of the current order (tets / edges: 4 / 2 mesh-local ids each) and,
 * with use_tris (the collision pass), the surface triangles (tris: int4
 * a, b, c, face id in BVH leaf order, like ffsrl_batch.tris) whose first
 * vertex it is the lowest part to hold.  Its vertices (local ids 0 ..) are
 * those its constraints in use and its triangles touch, in increasing mesh
 * id, plus the untouched vertices dealt round-robin in increasing id; verts
 * gets each part's vertices (mesh-local id | FFSRL_PART_SHARED when several
 * parts hold it | FFSRL_PART_OWNER on the lowest part holding it), rows the
 * parts' [8] rows (vertex and triangle entries relative to this mesh),
 * tets16 / edges16 each constraint's local ids and tris16 each part's
 * triangles' local ids, by part (zero for a family not in use), *n_verts
 * the entries written (verts_cap >= 4 nt + 2 ne + 3 nf + nv always
 * suffices) and *max_local the largest part.  No reference counterpart
 * (set-up of this library's partitioned frame). */
int ffsrl_partition_mesh(const int32_t* tets, int64_t nt, const int32_t* edges, int64_t ne, const int32_t* tris,
                         int64_t nf, int32_t nv, int32_t parts, int32_t use_tets, int32_t use_edges,
                         int32_t use_tris, int32_t* rows, int32_t* verts, int64_t verts_cap, int64_t* n_verts,
                         uint64_t* tets16, uint32_t* edges16, uint64_t* tris16, int32_t* max_local);

/* ---- HIP graphs --------------------------------------------------------- */
/* Capture everything issued on `stream` between begin and end into a graph
 * (replaces wp.capture_begin / capture_end / capture_launch). */
int ffsrl_graph_begin(void* stream);
int ffsrl_graph_end(void* stream, void** graph_exec);
int ffsrl_graph_launch(void* graph_exec, void* stream);
int ffsrl_graph_destroy(void* graph_exec);
/* Ends a capture begun with ffsrl_graph_begin and discards what it recorded
 * (for a caller whose work failed mid-capture; the stream leaves capture
 * mode).  No reference counterpart. */
int ffsrl_graph_abort(void* stream);

#ifdef __cplusplus
}
#endif
#endif
