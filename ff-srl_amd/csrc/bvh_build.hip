// Device binned-SAH BVH build with the exact topology of the host build
// (SimBVH.buildBVH / subdivide / findBestSplitPlaneBinned / updateNodeBounds,
// bvh.py:228-518; the host restatement is ffsrl_bvh_build in bvh.hip).
//
// One workgroup builds one mesh, top down, one tree level at a time; one
// wavefront processes one node of the level:
//   * centroid bounds per axis by wavefront min/max reductions;
//   * bin counts and bin boxes in LDS (float min/max on an order-preserving
//     integer encoding, so the result does not depend on the order);
//   * the bin sweep and the split choice by lane 0 in float64, operation by
//     operation as the host does;
//   * the reference's in-place two-pointer partition, computed in parallel
//     from its closed form (below);
//   * the children's bounds by wavefront reductions.
// Children get temporary ids in level order; a last pass gives every node the
// id the reference's depth-first recursion gives it (children pairs allocated
// in the order the recursion splits their parents).
//
// Partition closed form.  For a node of n triangles with predicate
// pred(p) = centroid < pos and L = #pred, the two-pointer loop leaves
//   * pred elements of [0, L) in place;
//   * the t-th (0-based) non-pred element of [0, L) ("hole" t) at g(t-1) - 1,
//     where g(t) is the position of the t-th pred element of [L, n) counted
//     from the right and g(-1) = n;
//   * pred element t (from the right) of [L, n) at hole t;
//   * a non-pred element of [L, n) at p - 1, except p = L, which goes to
//     g(r-1) - 1 (r = number of holes; g(r-1) = n when r = 0).
#include <cstdint>

#include "common.hpp"

namespace ffsrl {
namespace {

constexpr int kBuildBlock = 1024;
constexpr int kWaves = kBuildBlock / 64;
constexpr int kMaxBins = 32;
constexpr int kMaxDepth = 64;

// order-preserving float <-> uint32 (for LDS atomic min/max)
__device__ inline uint32_t enc(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float dec(uint32_t e) { return __uint_as_float((e & 0x80000000u) ? (e & 0x7fffffffu) : ~e); }

__device__ inline float wmin(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ inline float wmax(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

__device__ inline int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// ordering of this wavefront's LDS / global accesses across lanes
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

struct Job {
  const float* v;
  const int32_t* f;
  int nv, nf;
  // scratch
  float* cen;      // [nf][3]
  int32_t* tri;    // [nf] current permutation
  int32_t* tmp;    // [nf] partition output
  int32_t* holes;  // [nf]
  int32_t* gl;     // [nf] pred positions of [L, n), increasing
  int32_t* nfirst; int32_t* ncount; int32_t* nlevel; int32_t* nchild;  // [2 nf] per temporary node
  int32_t* nsize; int32_t* nid; int32_t* nbase;                       // [2 nf]
  float* nbox;     // [2 nf][6]
};

__device__ inline Job job_of(const ffsrl_bvh_job& j) {
  Job b;
  b.v = j.verts;
  b.f = j.faces;
  b.nv = j.nv;
  b.nf = j.nf;
  char* s = reinterpret_cast<char*>(j.scratch);
  const int64_t F = j.nf;
  b.cen = reinterpret_cast<float*>(s);
  s += 12 * F;
  b.tri = reinterpret_cast<int32_t*>(s);
  s += 4 * F;
  b.tmp = reinterpret_cast<int32_t*>(s);
  s += 4 * F;
  b.holes = reinterpret_cast<int32_t*>(s);
  s += 4 * F;
  b.gl = reinterpret_cast<int32_t*>(s);
  s += 4 * F;
  int32_t* n = reinterpret_cast<int32_t*>(s);
  b.nfirst = n;
  b.ncount = n + 2 * F;
  b.nlevel = n + 4 * F;
  b.nchild = n + 6 * F;
  b.nsize = n + 8 * F;
  b.nid = n + 10 * F;
  b.nbase = n + 12 * F;
  s += 4 * 14 * F;
  b.nbox = reinterpret_cast<float*>(s);
  return b;
}

__device__ inline const float* vtx(const Job& b, int t, int k) { return b.v + 3 * (int64_t)b.f[3 * (int64_t)t + k]; }

struct Bins {
  int cnt[kMaxBins];
  uint32_t lo[kMaxBins][3], hi[kMaxBins][3];
};

// Bounds of the triangles of [first, first + count) of the permutation.
__device__ inline void range_bounds(const Job& b, int first, int count, float* box, int lane) {
  float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
  for (int i = lane; i < count; i += 64) {
    const int t = b.tri[first + i];
    for (int k = 0; k < 3; ++k) {
      const float* p = vtx(b, t, k);
      for (int c = 0; c < 3; ++c) {
        mn[c] = fminf(mn[c], p[c]);
        mx[c] = fmaxf(mx[c], p[c]);
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    box[c] = wmin(mn[c]);
    box[3 + c] = wmax(mx[c]);
  }
}

// std::min / std::max of the host build
__device__ inline double dmin(double a, double b) { return b < a ? b : a; }
__device__ inline double dmax(double a, double b) { return a < b ? b : a; }

__device__ inline double area(const double* mn, const double* mx) {
  const double e0 = mx[0] - mn[0], e1 = mx[1] - mn[1], e2 = mx[2] - mn[2];
  return e0 * e1 + e1 * e2 + e2 * e0;
}

// findBestSplitPlaneBinned for node `node`: returns the cost, with axis / pos.
__device__ double best_split(const Job& b, int first, int count, int bins, Bins& sb, int lane, int* axis_out,
                             double* pos_out) {
  double best = 1e30, best_pos = 0.0;
  int best_axis = -1;
  for (int axis = 0; axis < 3; ++axis) {
    float cmn = 1e30f, cmx = -1e30f;
    for (int i = lane; i < count; i += 64) {
      const float c = b.cen[3 * b.tri[first + i] + axis];
      cmn = fminf(cmn, c);
      cmx = fmaxf(cmx, c);
    }
    cmn = wmin(cmn);
    cmx = wmax(cmx);
    const double bmin = cmn, bmax = cmx;
    if (bmax == bmin) continue;
    if (lane < bins) {
      sb.cnt[lane] = 0;
      for (int c = 0; c < 3; ++c) {
        sb.lo[lane][c] = enc(1e30f);
        sb.hi[lane][c] = enc(-1e30f);
      }
    }
    wave_sync();
    const double scale = (double)bins / (bmax - bmin);
    for (int i = lane; i < count; i += 64) {
      const int t = b.tri[first + i];
      const double c = b.cen[3 * t + axis];
      int id = (int)((c - bmin) * scale);
      if (id > bins - 1) id = bins - 1;
      atomicAdd(&sb.cnt[id], 1);
      const float *p0 = vtx(b, t, 0), *p1 = vtx(b, t, 1), *p2 = vtx(b, t, 2);
      for (int k = 0; k < 3; ++k) {
        atomicMin(&sb.lo[id][k], enc(fminf(fminf(p0[k], p1[k]), p2[k])));
        atomicMax(&sb.hi[id][k], enc(fmaxf(fmaxf(p0[k], p1[k]), p2[k])));
      }
    }
    wave_sync();
    if (lane == 0) {
      // the sweep of bvh.py:330-347 / Builder::best_split, in float64
      double lmn[3] = {1e30, 1e30, 1e30}, lmx[3] = {-1e30, -1e30, -1e30};
      double rmn[3] = {1e30, 1e30, 1e30}, rmx[3] = {-1e30, -1e30, -1e30};
      double la[kMaxBins], ra[kMaxBins];
      int lc[kMaxBins], rc[kMaxBins];
      int ls = 0, rs = 0;
      // an empty bin's box (+-1e30) changes no min / max, so it is skipped
      for (int j = 0; j < bins - 1; ++j) {
        ls += sb.cnt[j];
        lc[j] = ls;
        if (sb.cnt[j] > 0)
          for (int k = 0; k < 3; ++k) {
            lmn[k] = dmin(lmn[k], (double)dec(sb.lo[j][k]));
            lmx[k] = dmax(lmx[k], (double)dec(sb.hi[j][k]));
          }
        la[j] = area(lmn, lmx);
        const int jr = bins - 1 - j;
        rs += sb.cnt[jr];
        rc[bins - 2 - j] = rs;
        if (sb.cnt[jr] > 0)
          for (int k = 0; k < 3; ++k) {
            rmn[k] = dmin(rmn[k], (double)dec(sb.lo[jr][k]));
            rmx[k] = dmax(rmx[k], (double)dec(sb.hi[jr][k]));
          }
        ra[bins - 2 - j] = area(rmn, rmx);
      }
      const double step = (bmax - bmin) / bins;
      for (int j = 0; j < bins - 1; ++j) {
        const double cost = lc[j] * la[j] + rc[j] * ra[j];
        if (cost < best) {
          best_axis = axis;
          best_pos = bmin + step * (j + 1);
          best = cost;
        }
      }
    }
    best = __shfl(best, 0);
    best_pos = __shfl(best_pos, 0);
    best_axis = __shfl(best_axis, 0);
    wave_sync();  // lane 0 has read the bins before the next axis clears them
  }
  *axis_out = best_axis;
  *pos_out = best_pos;
  return best;
}

// The reference's in-place partition of [first, first + n) by
// centroid[axis] < pos, from its closed form; returns L.
__device__ int partition(const Job& b, int first, int n, int axis, double pos, int lane) {
  int L = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    const bool pr = i < n && (double)b.cen[3 * b.tri[first + i] + axis] < pos;
    L += __popcll(__ballot(pr));
  }
  // holes (non-pred of [0, L), increasing) and pred positions of [L, n) (increasing)
  int nh = 0, ng = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    const bool pr = i < n && (double)b.cen[3 * b.tri[first + i] + axis] < pos;
    const bool hole = i < L && !pr;
    const bool gp = i < n && i >= L && pr;
    const uint64_t mh = __ballot(hole), mg = __ballot(gp);
    if (hole) b.holes[first + nh + lanes_below(mh)] = i;
    if (gp) b.gl[first + ng + lanes_below(mg)] = i;
    nh += __popcll(mh);
    ng += __popcll(mg);
  }
  wave_sync();
  const int r = nh;  // == ng
  const int gr = r ? b.gl[first + 0] : n;
  int hseen = 0, gseen = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    const bool pr = i < n && (double)b.cen[3 * b.tri[first + i] + axis] < pos;
    const bool hole = i < L && !pr;
    const bool gp = i < n && i >= L && pr;
    const uint64_t mh = __ballot(hole), mg = __ballot(gp);
    if (i < n) {
      int dest;
      if (i < L) {
        if (pr) {
          dest = i;
        } else {
          const int t = hseen + lanes_below(mh);  // hole rank
          dest = (t == 0 ? n : b.gl[first + (r - t)]) - 1;
        }
      } else if (pr) {
        const int t = r - 1 - (gseen + lanes_below(mg));  // rank from the right
        dest = b.holes[first + t];
      } else {
        dest = (i == L) ? gr - 1 : i - 1;
      }
      b.tmp[first + dest] = b.tri[first + i];
    }
    hseen += __popcll(mh);
    gseen += __popcll(mg);
  }
  wave_sync();
  for (int i = lane; i < n; i += 64) b.tri[first + i] = b.tmp[first + i];
  wave_sync();
  return L;
}

__global__ __launch_bounds__(kBuildBlock) void k_bvh_build(const ffsrl_bvh_job* jobs, int bins) {
  const ffsrl_bvh_job& J = jobs[blockIdx.x];
  const Job b = job_of(J);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int F = b.nf;
  __shared__ Bins sbins[kWaves];
  __shared__ int lvl_start[kMaxDepth + 2];
  __shared__ int counter, max_level, overflow;
  __shared__ float rbox[6];

  // centroids (calculateCentroids: (a + b) + c, / 3 in float32) and the identity permutation
  for (int t = tid; t < F; t += kBuildBlock) {
    for (int k = 0; k < 3; ++k) {
      float s = vtx(b, t, 0)[k] + vtx(b, t, 1)[k];
      s = s + vtx(b, t, 2)[k];
      b.cen[3 * t + k] = s / 3.0f;
    }
    b.tri[t] = t;
  }
  // root bounds: all vertices
  if (wave == 0) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = lane; i < b.nv; i += 64)
      for (int c = 0; c < 3; ++c) {
        mn[c] = fminf(mn[c], b.v[3 * (int64_t)i + c]);
        mx[c] = fmaxf(mx[c], b.v[3 * (int64_t)i + c]);
      }
    for (int c = 0; c < 3; ++c) {
      const float lo = wmin(mn[c]), hi = wmax(mx[c]);
      if (lane == 0) {
        rbox[c] = lo;
        rbox[3 + c] = hi;
      }
    }
  }
  if (tid == 0) {
    counter = 1;
    max_level = 0;
    overflow = 0;
    lvl_start[0] = 0;
    lvl_start[1] = 1;
  }
  __syncthreads();
  if (tid == 0) {
    b.nfirst[0] = 0;
    b.ncount[0] = F;
    b.nlevel[0] = 0;
    b.nchild[0] = -1;
    for (int c = 0; c < 6; ++c) b.nbox[c] = rbox[c];
  }
  __syncthreads();

  int depth = 0;
  for (; depth < kMaxDepth; ++depth) {
    const int ls = lvl_start[depth], le = lvl_start[depth + 1];
    if (ls == le) break;
    for (int node = ls + wave; node < le; node += kWaves) {
      const int first = b.nfirst[node], count = b.ncount[node];
      int axis;
      double pos;
      const double best = best_split(b, first, count, bins, sbins[wave], lane, &axis, &pos);
      const float* nb = b.nbox + 6 * (int64_t)node;
      const double e0 = (double)nb[3] - (double)nb[0], e1 = (double)nb[4] - (double)nb[1],
                   e2 = (double)nb[5] - (double)nb[2];
      const double parent = count * (e0 * e1 + e1 * e2 + e2 * e0);
      if (best >= parent) continue;
      const int L = partition(b, first, count, axis, pos, lane);
      if (L == 0 || L == count) continue;
      int c = 0;
      if (lane == 0) c = atomicAdd(&counter, 2);
      c = __shfl(c, 0);
      if (c + 2 > 2 * F) {  // cannot happen for a binary tree over F leaves; stop safely
        if (lane == 0) overflow = 1;
        continue;
      }
      float bl[6], br[6];
      range_bounds(b, first, L, bl, lane);
      range_bounds(b, first + L, count - L, br, lane);
      if (lane == 0) {
        b.nfirst[c] = first;
        b.ncount[c] = L;
        b.nlevel[c] = depth + 1;
        b.nchild[c] = -1;
        b.nfirst[c + 1] = first + L;
        b.ncount[c + 1] = count - L;
        b.nlevel[c + 1] = depth + 1;
        b.nchild[c + 1] = -1;
        for (int k = 0; k < 6; ++k) {
          b.nbox[6 * (int64_t)c + k] = bl[k];
          b.nbox[6 * (int64_t)(c + 1) + k] = br[k];
        }
        b.nchild[node] = c;
        atomicMax(&max_level, depth + 1);
      }
    }
    __syncthreads();
    if (tid == 0) lvl_start[depth + 2] = counter;
    __syncthreads();
  }
  if (depth == kMaxDepth && tid == 0) overflow = 1;
  __syncthreads();
  const int used = counter;

  // subtree sizes, deepest level first
  for (int d = depth - 1; d >= 0; --d) {
    for (int v = lvl_start[d] + tid; v < lvl_start[d + 1]; v += kBuildBlock) {
      const int ch = b.nchild[v];
      b.nsize[v] = ch < 0 ? 1 : 1 + b.nsize[ch] + b.nsize[ch + 1];
    }
    __syncthreads();
  }
  // depth-first ids: a node's children pair is allocated when it splits, the
  // left subtree is finished before the right one (subdivide, bvh.py:349-401)
  if (tid == 0) {
    b.nid[0] = 0;
    b.nbase[0] = 1;
  }
  __syncthreads();
  for (int d = 0; d < depth; ++d) {
    for (int v = lvl_start[d] + tid; v < lvl_start[d + 1]; v += kBuildBlock) {
      const int ch = b.nchild[v];
      if (ch < 0) continue;
      const int base = b.nbase[v];
      b.nid[ch] = base;
      b.nid[ch + 1] = base + 1;
      b.nbase[ch] = base + 2;
      b.nbase[ch + 1] = base + 1 + b.nsize[ch];
    }
    __syncthreads();
  }
  // outputs in the host build's layout
  for (int i = tid; i < 2 * F; i += kBuildBlock) {
    J.left[i] = -1;
    J.first[i] = 0;
    J.count[i] = 0;
    J.level[i] = -1;
  }
  __syncthreads();
  for (int v = tid; v < used; v += kBuildBlock) {
    const int id = b.nid[v], ch = b.nchild[v];
    J.left[id] = ch < 0 ? -1 : b.nid[ch];
    J.first[id] = b.nfirst[v];
    J.count[id] = ch < 0 ? b.ncount[v] : 0;
    J.level[id] = b.nlevel[v];
  }
  for (int t = tid; t < F; t += kBuildBlock) J.tri_map[t] = b.tri[t];
  if (tid == 0) {
    J.info[0] = overflow ? -1 : used;
    J.info[1] = max_level;
  }
}

}  // namespace
}  // namespace ffsrl

using namespace ffsrl;

extern "C" {

int64_t ffsrl_bvh_scratch_bytes(int32_t nf) { return nf <= 0 ? 0 : (int64_t)nf * (12 + 16 + 4 * 14 + 48) + 256; }

int ffsrl_bvh_build_device(const ffsrl_bvh_job* jobs, int32_t n_jobs, int32_t bins, void* stream) {
  if (n_jobs < 0 || (n_jobs > 0 && !jobs)) return fail(FFSRL_E_ARG, "bvh_build_device: no jobs");
  if (bins < 2 || bins > kMaxBins) return fail(FFSRL_E_ARG, "bvh_build_device: bins %d outside [2, %d]", bins, kMaxBins);
  if (n_jobs == 0) return 0;
  hipLaunchKernelGGL(k_bvh_build, dim3(n_jobs), dim3(kBuildBlock), 0, as_stream(stream), jobs, bins);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
