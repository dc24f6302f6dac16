// Ray-cast renderer: drawSimBVHKernelManyEnvs + meshQueryRayManyEnvs +
// downsample_kernel (render.py:276-498, 699-724) and the vertex normals
// (render.py:81-121).
//
// One wavefront renders an 8x8-pixel tile of one environment; each lane owns
// one pixel, its quad of 2x-supersampled rays and a 32-entry traversal stack
// in LDS (slot-major, so the 64 lanes of a slot fall in distinct banks).  The
// 2x2 box filter sums the quad in the lane in the reference's summation
// order, so no ray buffer goes to HBM.  Traversal order, the Moller-Trumbore
// test and the strict "0 < t < tMin" acceptance follow the reference
// expression by expression for every ray, and the box tests' decisions equal
// the reference's slab test (a fast form with an exact fallback,
// render_box.hpp), so hit triangles are bit-exact.
#include <algorithm>
#include <climits>

#include "common.hpp"
#include "render_box.hpp"

namespace ffsrl {

constexpr int kStack = 32;  // deepest supported BVH (levels); the reference's stack size
constexpr int kWaves = 1;   // wavefronts per block (one 8x8 ray tile each)
constexpr int kXcds = 8;    // gfx950 dispatches workgroups round-robin over 8 XCDs

__device__ __forceinline__ float mt_isect(f3 a, f3 b, f3 c, f3 o, f3 d) {
  const f3 e1 = b - a, e2 = c - a;
  const f3 pv = cross(d, e2);
  const float det = dot(e1, pv);
  if (fabsf(det) < 0.0001f) return -1.0f;
  const float inv = 1.0f / det;
  const f3 tv = o - a;
  const float u = dot(tv, pv) * inv;
  if (u < 0.0f || u > 1.0f) return -1.0f;
  const f3 qv = cross(tv, e1);
  const float v = dot(d, qv) * inv;
  if (v < 0.0f || u + v > 1.0f) return -1.0f;
  return dot(e2, qv) * inv;
}

template <typename T>
__device__ __forceinline__ T ldu4(const T* base, int i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (uint32_t)i * (uint32_t)sizeof(T));
}

// float -> int as the reference's int() on the device: truncation toward
// zero, NaN to 0, saturating at the int range.
__device__ __forceinline__ int trunc_i(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483647.0f) return 2147483647;
  if (x <= -2147483648.0f) return -2147483647 - 1;
  return (int)x;
}

// One texel of the concatenated table; an index outside it reads as zero
// (the reference's lookup reads one column past a texture's last, where the
// weight is zero).
__device__ __forceinline__ f3 texel(const ffsrl_batch& b, int64_t i) {
  if (i < 0 || i >= b.n_texels) return mk(0.0f, 0.0f, 0.0f);
  return xyz(ld4(b.texels, i));
}

// texture_interpolation (render.py:124-148): bilinear lookup in square
// texture tx at texture coordinates (s, t), with the reference's indexing
// y + x * width and no clamp of x0 + 1 / y0 + 1.
__device__ __forceinline__ f3 texture_lookup(const ffsrl_batch& b, int tx, float s, float t) {
  const int64_t first = b.tex_info[2 * tx];
  const int size = (int)b.tex_info[2 * tx + 1];
  const float fx = s * (float)(size - 1), fy = (1.0f - t) * (float)(size - 1);
  const int x0 = trunc_i(fx), y0 = trunc_i(fy);
  const int x1 = x0 + 1, y1 = y0 + 1;
  const float ax = fx - (float)x0, ay = fy - (float)y0;
  const f3 c00 = texel(b, first + y0 + (int64_t)x0 * size), c10 = texel(b, first + y0 + (int64_t)x1 * size);
  const f3 c01 = texel(b, first + y1 + (int64_t)x0 * size), c11 = texel(b, first + y1 + (int64_t)x1 * size);
  const f3 lower = c00 * (1.0f - ax) + c10 * ax;
  const f3 upper = c01 * (1.0f - ax) + c11 * ax;
  return lower * (1.0f - ay) + upper * ay;
}

struct RenderOut {
  float* depth;
  float* rgb;
  const float* normals;
  const float* vcolor;
  const float* tcolor;
  int32_t* hit;
  float* thit;
};

// Shading of one ray's hit (colour modes; mode 4 gives depth in .x):
// render.py:356-409, expression by expression.
__device__ __forceinline__ f3 shade(const ffsrl_batch& b, const ffsrl_camera& cam, const RenderOut& out, int64_t vb,
                                    int64_t fb, f3 o, f3 d, float tMin, int best, int4 tri) {
  f3 color;
  if (cam.mode == 4) {
    const float c = tMin / 50.0f * 1.0f;
    color = mk(c, c, c);
  } else {
    const f3 A = xyz(ld4(b.x, vb + tri.x)), B = xyz(ld4(b.x, vb + tri.y)), C = xyz(ld4(b.x, vb + tri.z));
    const f3 p = o + d * tMin;
    const float area = length(cross(B - A, C - A));
    const float w = length(cross(B - A, p - A)) / area;
    const float v = length(cross(p - A, C - A)) / area;
    const float u = 1.0f - w - v;
    const f3 na = xyz(ld4(out.normals, vb + tri.x)), nb = xyz(ld4(out.normals, vb + tri.y)),
             nc = xyz(ld4(out.normals, vb + tri.z));
    const f3 normal = (na * u + nb * v) + nc * w;
    if (cam.mode == 2) {
      color = normal * 0.5f + mk(0.5f, 0.5f, 0.5f);
    } else {
      const int tex = (cam.mode == 1 && b.tri_tex) ? b.tri_tex[fb + best] : -1;
      if (cam.mode == 0) {
        color = mk(1.0f, 1.0f, 1.0f);
      } else if (cam.mode == 5) {
        const float* tc = out.tcolor + 3 * (int64_t)(fb + best);
        color = mk(tc[0], tc[1], tc[2]);
      } else if (tex >= 0) {  // texture mode: tex = u * tex_a + v * tex_b + w * tex_c (render.py:356-377)
        const float* uv = b.tri_uv + 6 * (int64_t)(fb + best);
        const float s = (u * uv[0] + v * uv[2]) + w * uv[4];
        const float t = (u * uv[1] + v * uv[3]) + w * uv[5];
        color = texture_lookup(b, tex, s, t);
      } else {  // vertex colour, and the texture mode's fallback (render.py:378-380)
        const float* ca = out.vcolor + 3 * (vb + tri.x);
        const float* cb = out.vcolor + 3 * (vb + tri.y);
        const float* cc = out.vcolor + 3 * (vb + tri.z);
        color = ((mk(ca[0], ca[1], ca[2]) + mk(cb[0], cb[1], cb[2])) + mk(cc[0], cc[1], cc[2])) / 3.0f;
      }
      float val = cam.light_intensity * dot(normal, mk(cam.light_dir[0], cam.light_dir[1], cam.light_dir[2]));
      if (val < 0.0f) val = 0.0f;
      const float lambert = 0.0f + val;
      color = color * lambert;
    }
  }
  if (color.x > 1.0f) color.x = 1.0f;
  if (color.y > 1.0f) color.y = 1.0f;
  if (color.z > 1.0f) color.z = 1.0f;
  return color;
}

// A stack entry is a node id (< 2^28) and, in its top 4 bits, the rays of
// the lane's quad whose own traversal reaches the node (they hit its parent).
constexpr int kQuadShift = 28;
constexpr int kNodeMask = (1 << kQuadShift) - 1;
// Automatic choice: the quad traversal from 128x128 output pixels.  Measured
// on the MI355X against the per-ray form: at 64x64 depth 7% slower on mixed
// 1k-40k-tet envs and 35% slower on the 545-vertex liver scene (neighbouring
// rays of a quad part early in a coarse grid, and the quad loop's mask work
// is paid for rays that have left); at 128x128 depth even on the mixed envs,
// 3.5% faster on the liver scene, and 15% faster for 128x128 RGB-D on
// 20k-tet ellipsoids.
constexpr int kQuadMinPixels = 128 * 128;

// Environments: env_begin + el for el < n_env, or list[el] when list is not
// NULL; outputs are indexed by el for a range and by the env id for a list.
//
// One lane renders one pixel: its quad of supersampled rays k = 0..3, ray
// (2 py + (k >> 1), 2 px + (k & 1)).  The quad shares one traversal stack
// whose entries carry the mask of quad rays that reach the node: a node is
// tested for exactly those rays, and its children are pushed with the rays
// that hit it.  Every ray therefore tests the nodes of its own right-first
// DFS, in the same order and against its own tMin, so its hits are those of
// a traversal of its own; the quad shares the node and triangle loads, the
// stack and the loop.  A wavefront takes an 8x8-pixel tile of one env.
__global__ __launch_bounds__(64 * kWaves) void k_render_quad(ffsrl_batch b, ffsrl_camera cam, int env_begin,
                                                             const int32_t* list, int n_env, RenderOut out,
                                                             int depth) {
  extern __shared__ int stack[];  // [kWaves][depth][64]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int RW = 2 * cam.width;
  const int tiles_x = (cam.width + 7) >> 3, tiles = tiles_x * ((cam.height + 7) >> 3);
  // XCD-aware mapping of the 1D grid: block B runs on XCD B % 8, so the
  // blocks of env el are B = 8 * (bpe * (el / 8) + j) + el % 8, j < bpe: all
  // tiles of an env share one XCD's L2 (its BVH, triangles and positions are
  // fetched into one L2 instead of eight).
  const int bpe = (tiles + kWaves - 1) / kWaves;
  const int xcd = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
  const int el = (k / bpe) * kXcds + xcd;
  const int tile = (k % bpe) * kWaves + wave;
  if (el >= n_env || tile >= tiles) return;  // whole wavefront exits; no block barrier below
  const int px = (tile % tiles_x) * 8 + (lane & 7), py = (tile / tiles_x) * 8 + (lane >> 3);
  if (px >= cam.width || py >= cam.height) return;  // no cross-lane operation below
  const int env = list ? list[el] : env_begin + el;
  const int64_t oi = list ? env : el;  // output slot
  const ffsrl_env_desc e = b.envs[env];
  const int64_t vb = e.vbase, sb = e.boundbase, fb = e.fbase;
  const f3 o = mk(cam.origin[0], cam.origin[1], cam.origin[2]);
  f3 d[4], inv[4];
  bool ray_ok[4];
  float tMin[4];
  int best[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = (2 * py + (q >> 1)) * RW + 2 * px + (q & 1);
    d[q] = xyz(ld4(cam.dirs, r));
    inv[q] = mk(1.0f / d[q].x, 1.0f / d[q].y, 1.0f / d[q].z);
    ray_ok[q] = ray_dir_ok(d[q]);
    tMin[q] = 1e6f;
    best[q] = -1;
  }

  // env-uniform bases; node / triangle / vertex ids are non-negative int32,
  // so loads take a 32-bit unsigned offset from a scalar base
  const float4* blo = reinterpret_cast<const float4*>(b.bvh_lo) + sb;
  const float4* bhi = reinterpret_cast<const float4*>(b.bvh_hi) + sb;
  const int4* tris = reinterpret_cast<const int4*>(b.tris) + fb;
  const float4* X = reinterpret_cast<const float4*>(b.x) + vb;
  int* stk = stack + wave * depth * 64 + lane;
  stk[0] = 0xF << kQuadShift;  // the root, reached by all four rays
  int count = 1;
  while (count) {
    const int entry = stk[(count - 1) * 64];
    count -= 1;
    const int node = entry & kNodeMask, reach = (int)((unsigned)entry >> kQuadShift);
    const float4 lo = ldu4(blo, node), hi = ldu4(bhi, node);
    int hits = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (((reach >> q) & 1) && visit_node(o, d[q], inv[q], xyz(lo), xyz(hi), tMin[q], ray_ok[q])) hits |= 1 << q;
    if (!hits) continue;
    const int n = __float_as_int(hi.w);
    if (n > 0) {
      const int first = __float_as_int(lo.w);
      for (int k2 = 0; k2 < n; ++k2) {
        const int4 t = ldu4(tris, first + k2);
        const f3 A = xyz(ldu4(X, t.x)), B = xyz(ldu4(X, t.y)), C = xyz(ldu4(X, t.z));
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!((hits >> q) & 1)) continue;
          const float tt = mt_isect(A, B, C, o, d[q]);
          if (tt < tMin[q] && tt > 0.0f) {
            tMin[q] = tt;
            best[q] = first + k2;
          }
        }
      }
    } else if (count + 2 <= depth) {  // depth = deepest BVH level + 1: never false
      const int l = __float_as_int(lo.w);
      stk[count * 64] = l | (hits << kQuadShift);
      stk[(count + 1) * 64] = (l + 1) | (hits << kQuadShift);
      count += 2;
    }
  }

  const int64_t rays_env = (int64_t)RW * 2 * cam.height;
  const int64_t pix = oi * cam.height * cam.width + (int64_t)py * cam.width + px;
  f3 color[4];
  float dz[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool hitp = tMin[q] < 1e6f;
    int4 tri = make_int4(0, 0, 0, -1);
    if (hitp) tri = ldi4(b.tris, fb + best[q]);
    const int64_t r = (int64_t)(2 * py + (q >> 1)) * RW + 2 * px + (q & 1);
    if (out.hit) out.hit[oi * rays_env + r] = hitp ? tri.w : -1;
    if (out.thit) out.thit[oi * rays_env + r] = hitp ? tMin[q] : 0.0f;
    color[q] = hitp ? shade(b, cam, out, vb, fb, o, d[q], tMin[q], best[q], tri) : mk(0.0f, 0.0f, 0.0f);
    dz[q] = 0.0f;
    if (hitp) {
      dz[q] = tMin[q] / 50.0f * 1.0f;
      if (dz[q] > 1.0f) dz[q] = 1.0f;
    }
  }
  // 2x2 box: rays (2py+i, 2px+j) summed in (i, j) = (0,0), (0,1), (1,0), (1,1) order.
  if (cam.mode == 4) {
    if (out.depth) out.depth[pix] = (((color[0].x + color[1].x) + color[2].x) + color[3].x) / 4.0f;
  } else {
    if (out.rgb) {
      out.rgb[3 * pix + 0] = (((color[0].x + color[1].x) + color[2].x) + color[3].x) / 4.0f;
      out.rgb[3 * pix + 1] = (((color[0].y + color[1].y) + color[2].y) + color[3].y) / 4.0f;
      out.rgb[3 * pix + 2] = (((color[0].z + color[1].z) + color[2].z) + color[3].z) / 4.0f;
    }
    // RGB-D: the depth image of the same hits, computed as depth mode does
    if (out.depth) out.depth[pix] = (((dz[0] + dz[1]) + dz[2]) + dz[3]) / 4.0f;
  }
}

// Environments: env_begin + el for el < n_env, or list[el] when list is not
// NULL; outputs are indexed by el for a range and by the env id for a list.
//
// Per-ray form: one lane per supersampled ray, a wavefront per 8x8-ray tile;
// the 2x2 box filter sums across lanes with shuffles.
__global__ __launch_bounds__(64 * kWaves) void k_render(ffsrl_batch b, ffsrl_camera cam, int env_begin,
                                                        const int32_t* list, int n_env, RenderOut out, int depth) {
  extern __shared__ int stack[];  // [kWaves][depth][64]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int RW = 2 * cam.width, RH = 2 * cam.height;
  const int tiles_x = RW >> 3, tiles = tiles_x * (RH >> 3);
  // XCD-aware mapping of the 1D grid: block B runs on XCD B % 8, so the
  // blocks of env el are B = 8 * (bpe * (el / 8) + j) + el % 8, j < bpe: all
  // tiles of an env share one XCD's L2 (its BVH, triangles and positions are
  // fetched into one L2 instead of eight).
  const int bpe = (tiles + kWaves - 1) / kWaves;
  const int xcd = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
  const int el = (k / bpe) * kXcds + xcd;
  const int tile = (k % bpe) * kWaves + wave;
  if (el >= n_env || tile >= tiles) return;  // whole wavefront exits; no block barrier below
  const int env = list ? list[el] : env_begin + el;
  const int64_t oi = list ? env : el;  // output slot
  const int col = (tile % tiles_x) * 8 + (lane & 7), row = (tile / tiles_x) * 8 + (lane >> 3);
  const int r = row * RW + col;
  const ffsrl_env_desc e = b.envs[env];
  const int64_t vb = e.vbase, sb = e.boundbase, fb = e.fbase;
  const f3 o = mk(cam.origin[0], cam.origin[1], cam.origin[2]);
  const f3 d = xyz(ld4(cam.dirs, r));
  const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  const bool ray_ok = ray_dir_ok(d);

  // env-uniform bases; node / triangle / vertex ids are non-negative int32,
  // so loads take a 32-bit unsigned offset from a scalar base
  const float4* blo = reinterpret_cast<const float4*>(b.bvh_lo) + sb;
  const float4* bhi = reinterpret_cast<const float4*>(b.bvh_hi) + sb;
  const int4* tris = reinterpret_cast<const int4*>(b.tris) + fb;
  const float4* X = reinterpret_cast<const float4*>(b.x) + vb;
  int* stk = stack + wave * depth * 64 + lane;
  stk[0] = 0;
  int count = 1, best = -1;
  float tMin = 1e6f;
  while (count) {
    const int node = stk[(count - 1) * 64];
    count -= 1;
    const float4 lo = ldu4(blo, node), hi = ldu4(bhi, node);
    if (visit_node(o, d, inv, xyz(lo), xyz(hi), tMin, ray_ok)) {
      const int n = __float_as_int(hi.w);
      if (n > 0) {
        const int first = __float_as_int(lo.w);
        for (int k = 0; k < n; ++k) {
          const int4 t = ldu4(tris, first + k);
          const float tt = mt_isect(xyz(ldu4(X, t.x)), xyz(ldu4(X, t.y)), xyz(ldu4(X, t.z)), o, d);
          if (tt < tMin && tt > 0.0f) {
            tMin = tt;
            best = first + k;
          }
        }
      } else if (count + 2 <= depth) {  // depth = deepest BVH level + 1: never false
        const int l = __float_as_int(lo.w);
        stk[count * 64] = l;
        stk[(count + 1) * 64] = l + 1;
        count += 2;
      }
    }
  }
  const bool hitp = tMin < 1e6f;
  int4 tri = make_int4(0, 0, 0, -1);
  if (hitp) tri = ldi4(b.tris, fb + best);
  const int64_t rays_env = (int64_t)RW * RH;
  if (out.hit) out.hit[oi * rays_env + r] = hitp ? tri.w : -1;
  if (out.thit) out.thit[oi * rays_env + r] = hitp ? tMin : 0.0f;

  const f3 color = hitp ? shade(b, cam, out, vb, fb, o, d, tMin, best, tri) : mk(0.0f, 0.0f, 0.0f);
  // 2x2 box: rays (2py+i, 2px+j) summed in (i, j) = (0,0), (0,1), (1,0), (1,1) order.
  const bool owner = ((lane & 1) == 0) && (((lane >> 3) & 1) == 0);
  const int64_t pix = oi * cam.height * cam.width + (row >> 1) * cam.width + (col >> 1);
  if (cam.mode == 4) {
    const float c01 = __shfl_down(color.x, 1), c10 = __shfl_down(color.x, 8), c11 = __shfl_down(color.x, 9);
    if (owner && out.depth) out.depth[pix] = (((color.x + c01) + c10) + c11) / 4.0f;
  } else {
    float ch[3] = {color.x, color.y, color.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float c01 = __shfl_down(ch[k], 1), c10 = __shfl_down(ch[k], 8), c11 = __shfl_down(ch[k], 9);
      if (owner && out.rgb) out.rgb[3 * pix + k] = (((ch[k] + c01) + c10) + c11) / 4.0f;
    }
    if (out.depth) {
      // RGB-D: the depth image of the same hits, computed as depth mode does
      float dz = 0.0f;
      if (hitp) {
        dz = tMin / 50.0f * 1.0f;
        if (dz > 1.0f) dz = 1.0f;
      }
      const float d01 = __shfl_down(dz, 1), d10 = __shfl_down(dz, 8), d11 = __shfl_down(dz, 9);
      if (owner) out.depth[pix] = (((dz + d01) + d10) + d11) / 4.0f;
    }
  }
}

// Area-weighted vertex normals: one thread per vertex sums the area normals
// cross(b - a, c - a) of its faces in increasing face id, then normalizes.
__global__ __launch_bounds__(256) void k_normals(ffsrl_batch b, int env_begin, float* normals) {
  const ffsrl_env_desc e = b.envs[env_begin + blockIdx.y];
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= e.nv) return;
  const int64_t vb = e.vbase;
  const int32_t* off = b.vf_off + e.vtopo_base;
  const int32_t* idx = b.vf_idx + 3 * (int64_t)e.fbase;
  f3 s = mk(0.0f, 0.0f, 0.0f);
  for (int k = off[v]; k < off[v + 1]; ++k) {
    const int4 t = ldi4(b.tris, (int64_t)e.fbase + idx[k]);
    const f3 A = xyz(ld4(b.x, vb + t.x)), B = xyz(ld4(b.x, vb + t.y)), C = xyz(ld4(b.x, vb + t.z));
    s = s + cross(B - A, C - A);
  }
  const f3 n = normalize(s);
  st4(normals, vb + v, make_float4(n.x, n.y, n.z, 0.0f));
}

}  // namespace ffsrl

using namespace ffsrl;

extern "C" {

static int render(const ffsrl_batch* b, const ffsrl_camera* cam, int32_t eb, const int32_t* list, int32_t n,
                  const RenderOut& out, hipStream_t st) {
  if (!b || !cam || !b->envs || !b->x || !b->tris || !b->bvh_lo || !b->bvh_hi || !cam->dirs)
    return fail(FFSRL_E_ARG, "render needs positions, triangles, BVH bounds and ray directions");
  if (cam->width <= 0 || cam->height <= 0 || (cam->width & 3) || (cam->height & 3))
    return fail(FFSRL_E_ARG, "render size %dx%d: width and height must be positive multiples of 4", cam->height,
                cam->width);
  const int m = cam->mode;
  if (m < 0 || m > 5) return fail(FFSRL_E_ARG, "render mode %d", m);
  if (m == 4 && !out.depth && !out.hit) return fail(FFSRL_E_ARG, "depth mode needs a depth (or hit) output");
  // a colour mode also writes the depth image of the same traversal when a
  // depth output is given (RGB-D in one pass)
  if (m != 4 && (!out.rgb || !out.normals)) return fail(FFSRL_E_ARG, "colour modes need rgb output and normals");
  if ((m == 3 || m == 1) && !out.vcolor) return fail(FFSRL_E_ARG, "vertex colour and texture modes need vertex colours");
  if (m == 1 && b->tri_tex && (!b->tri_uv || !b->texels || !b->tex_info))
    return fail(FFSRL_E_ARG, "texture mode: triangle textures given without coordinates, texels or texture table");
  if (m == 5 && !out.tcolor) return fail(FFSRL_E_ARG, "triangle colour mode needs triangle colours");
  if (cam->traversal < 0 || cam->traversal > 2) return fail(FFSRL_E_ARG, "render traversal %d", cam->traversal);
  if (n == 0) return 0;
  const bool quad = cam->traversal == FFSRL_TRAVERSE_QUAD ||
                    (cam->traversal == FFSRL_TRAVERSE_AUTO && cam->width * cam->height >= kQuadMinPixels);
  const int tiles = quad ? ((cam->width + 7) / 8) * ((cam->height + 7) / 8) : (cam->width / 4) * (cam->height / 4);
  // Stack entries per ray: the deepest BVH of the envs (levels = max level + 1
  // bounds the stack of this right-first DFS; a list is bounded by the whole
  // table); kStack without a host mirror.
  int stack_depth = kStack;
  if (b->host_envs) {
    stack_depth = 2;
    const int lo = list ? 0 : eb, hi = list ? b->n_env : eb + n;
    for (int i = lo; i < hi; ++i) {
      stack_depth = std::max(stack_depth, (int)b->host_envs[i].bvh_depth + 1);
      if (quad && b->host_envs[i].nn > kNodeMask)  // a quad stack entry holds a node id in 28 bits
        return fail(FFSRL_E_LIMIT, "env %d: %d BVH nodes exceed the ray stack's %d", i, b->host_envs[i].nn, kNodeMask);
    }
    if (stack_depth > kStack) return fail(FFSRL_E_LIMIT, "BVH with %d levels exceeds the %d-entry ray stack",
                                          stack_depth - 1, kStack);
  }
  const int lds = kWaves * stack_depth * 64 * (int)sizeof(int);
  const int64_t bpe = (tiles + kWaves - 1) / kWaves, groups = (n + kXcds - 1) / kXcds;
  if (groups * kXcds * bpe > INT32_MAX) return fail(FFSRL_E_LIMIT, "render grid too large");
  if (quad)
    hipLaunchKernelGGL(k_render_quad, dim3((unsigned)(groups * kXcds * bpe)), dim3(64 * kWaves), lds, st, *b, *cam,
                       eb, list, n, out, stack_depth);
  else
    hipLaunchKernelGGL(k_render, dim3((unsigned)(groups * kXcds * bpe)), dim3(64 * kWaves), lds, st, *b, *cam, eb,
                       list, n, out, stack_depth);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_render(const ffsrl_batch* b, const ffsrl_camera* cam, int32_t eb, int32_t ee, float* depth, float* rgb,
                 const float* normals, const float* vcolor, const float* tcolor, int32_t* hit, float* thit,
                 void* stream) {
  if (!b) return fail(FFSRL_E_ARG, "batch NULL");
  if (eb < 0 || ee > b->n_env || eb > ee) return fail(FFSRL_E_ARG, "env range [%d,%d) outside [0,%d)", eb, ee, b->n_env);
  return render(b, cam, eb, nullptr, ee - eb, RenderOut{depth, rgb, normals, vcolor, tcolor, hit, thit},
                as_stream(stream));
}

int ffsrl_render_envs(const ffsrl_batch* b, const ffsrl_camera* cam, const int32_t* envs, int32_t n, float* depth,
                      float* rgb, const float* normals, const float* vcolor, const float* tcolor, int32_t* hit,
                      float* thit, void* stream) {
  if (!b) return fail(FFSRL_E_ARG, "batch NULL");
  if (n < 0 || n > b->n_env || (n > 0 && !envs))
    return fail(FFSRL_E_ARG, "env list of %d entries (batch has %d envs)", n, b->n_env);
  return render(b, cam, 0, envs, n, RenderOut{depth, rgb, normals, vcolor, tcolor, hit, thit}, as_stream(stream));
}

int ffsrl_vertex_normals(const ffsrl_batch* b, int32_t eb, int32_t ee, float* normals, void* stream) {
  if (!b || !b->envs || !b->x || !b->tris || !normals || !b->vf_off || !b->vf_idx)
    return fail(FFSRL_E_ARG, "normals: NULL argument (positions, triangles, vertex-face adjacency, output)");
  if (eb < 0 || ee > b->n_env || eb > ee) return fail(FFSRL_E_ARG, "env range [%d,%d) outside [0,%d)", eb, ee, b->n_env);
  if (ee == eb) return 0;
  if (ee - eb > 65535) return fail(FFSRL_E_LIMIT, "at most 65535 environments per normals call");
  const int nv = b->max_nv < 1 ? 1 : b->max_nv;
  hipLaunchKernelGGL(k_normals, dim3((nv + 255) / 256, ee - eb), dim3(256), 0, as_stream(stream), *b, eb, normals);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
