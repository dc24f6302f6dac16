// BVH: refit kernels (batched and plain) and the host binned-SAH build.
#include <algorithm>
#include <cmath>
#include <vector>

#include "refit.hpp"

namespace ffsrl {

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_refit(ffsrl_batch b, int env_begin) {
  refit_env<BLOCK>(b, b.envs[env_begin + blockIdx.x]);
}

__global__ void k_refit_level(ffsrl_batch b, ffsrl_env_desc e, int beg, int end) {
  const int k = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (k < end) refit_node(b, e, k);
}

// ------------------------------------------------------------- host build
//
// SimBVH.buildBVH / subdivide / findBestSplitPlaneBinned / updateNodeBounds
// (bvh.py:281-401, 456-518).  Centroids are (a+b+c)/3 in float32
// (calculateCentroids, bvh.py:6-21); bin edges, areas and costs are float64
// as in the reference's NumPy code; the root bounds are those of all
// vertices; children are allocated in pairs after the root and the
// partition is the reference's in-place two-pointer swap.
namespace {

struct Box {
  double mn[3] = {1e30, 1e30, 1e30};
  double mx[3] = {-1e30, -1e30, -1e30};
  void grow(const float* p) {
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::min(mn[k], (double)p[k]);
      mx[k] = std::max(mx[k], (double)p[k]);
    }
  }
  void grow(const Box& o) {
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::min(mn[k], o.mn[k]);
      mx[k] = std::max(mx[k], o.mx[k]);
    }
  }
  double area() const {
    double e0 = mx[0] - mn[0], e1 = mx[1] - mn[1], e2 = mx[2] - mn[2];
    return e0 * e1 + e1 * e2 + e2 * e0;
  }
};

struct Builder {
  const float* v;
  const int32_t* f;
  int bins;
  std::vector<float> cen;  // [nf][3]
  std::vector<int32_t> tri;
  std::vector<Box> box;
  int32_t *left, *first, *count, *level;
  int used = 1;
  int max_level = 0;

  const float* vert(int t, int k) const { return v + 3 * (int64_t)f[3 * (int64_t)t + k]; }

  void bounds(int node) {
    Box b;
    for (int i = first[node]; i < first[node] + count[node]; ++i)
      for (int k = 0; k < 3; ++k) b.grow(vert(tri[i], k));
    box[node] = b;
  }

  double best_split(int node, int* axis_out, double* pos_out) {
    double best = 1e30;
    int best_axis = -1;
    double best_pos = 0.0;
    const int fst = first[node], cnt = count[node];
    std::vector<Box> bb(bins);
    std::vector<int> bc(bins);
    std::vector<double> la(bins - 1), ra(bins - 1);
    std::vector<int> lc(bins - 1), rc(bins - 1);
    for (int axis = 0; axis < 3; ++axis) {
      double bmin = 1e30, bmax = -1e30;
      for (int j = 0; j < cnt; ++j) {
        double c = cen[3 * tri[fst + j] + axis];
        bmin = std::min(bmin, c);
        bmax = std::max(bmax, c);
      }
      if (bmax == bmin) continue;
      std::fill(bb.begin(), bb.end(), Box());
      std::fill(bc.begin(), bc.end(), 0);
      double scale = (double)bins / (bmax - bmin);
      for (int j = 0; j < cnt; ++j) {
        const int t = tri[fst + j];
        double c = cen[3 * t + axis];
        int id = std::min(bins - 1, (int)((c - bmin) * scale));
        bc[id] += 1;
        for (int k = 0; k < 3; ++k) bb[id].grow(vert(t, k));
      }
      Box lb, rb;
      int ls = 0, rs = 0;
      for (int j = 0; j < bins - 1; ++j) {
        ls += bc[j];
        lc[j] = ls;
        lb.grow(bb[j]);
        la[j] = lb.area();
        rs += bc[bins - 1 - j];
        rc[bins - 2 - j] = rs;
        rb.grow(bb[bins - 1 - j]);
        ra[bins - 2 - j] = rb.area();
      }
      double step = (bmax - bmin) / bins;
      for (int j = 0; j < bins - 1; ++j) {
        double cost = lc[j] * la[j] + rc[j] * ra[j];
        if (cost < best) {
          best_axis = axis;
          best_pos = bmin + step * (j + 1);
          best = cost;
        }
      }
    }
    *axis_out = best_axis;
    *pos_out = best_pos;
    return best;
  }

  void subdivide(int node, int lvl) {
    int axis;
    double pos;
    double best = best_split(node, &axis, &pos);
    const Box& nb = box[node];
    double e0 = nb.mx[0] - nb.mn[0], e1 = nb.mx[1] - nb.mn[1], e2 = nb.mx[2] - nb.mn[2];
    double parent = count[node] * (e0 * e1 + e1 * e2 + e2 * e0);
    if (best >= parent) return;
    int i = first[node], j = i + count[node] - 1;
    while (i <= j) {
      if ((double)cen[3 * tri[i] + axis] < pos) {
        ++i;
      } else {
        std::swap(tri[i], tri[j]);
        --j;
      }
    }
    int lcount = i - first[node];
    if (lcount == 0 || lcount == count[node]) return;
    int lc = used, rc = used + 1;
    used += 2;
    first[lc] = first[node];
    count[lc] = lcount;
    level[lc] = lvl + 1;
    first[rc] = i;
    count[rc] = count[node] - lcount;
    level[rc] = lvl + 1;
    left[node] = lc;
    count[node] = 0;
    max_level = std::max(max_level, lvl + 1);
    bounds(lc);
    bounds(rc);
    subdivide(lc, lvl + 1);
    subdivide(rc, lvl + 1);
  }
};

}  // namespace
}  // namespace ffsrl

using namespace ffsrl;

extern "C" {

int ffsrl_bvh_refit(const ffsrl_batch* b, int32_t eb, int32_t ee, void* stream) {
  if (!b || !b->envs || !b->x || !b->tris || !b->node_left || !b->node_first || !b->node_count ||
      !b->level_nodes || !b->level_info || !b->level_off || !b->bvh_lo || !b->bvh_hi)
    return fail(FFSRL_E_ARG, "refit needs positions, triangles, BVH topology and bound arrays");
  if (eb < 0 || ee > b->n_env || eb > ee) return fail(FFSRL_E_ARG, "env range [%d,%d) outside [0,%d)", eb, ee, b->n_env);
  if (ee == eb) return 0;
  hipLaunchKernelGGL(k_refit<256>, dim3(ee - eb), dim3(256), 0, as_stream(stream), *b, eb);
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_bvh_refit_plain(const ffsrl_batch* b, int32_t eb, int32_t ee, void* stream) {
  if (!b || !b->envs || !b->level_off || !b->level_nodes || !b->level_info || !b->bvh_lo || !b->bvh_hi || !b->tris)
    return fail(FFSRL_E_ARG, "refit needs the env table and BVH arrays");
  if (eb < 0 || ee > b->n_env || eb > ee) return fail(FFSRL_E_ARG, "env range [%d,%d) outside [0,%d)", eb, ee, b->n_env);
  hipStream_t st = as_stream(stream);
  // The per-level launches are issued from the host: use the host mirrors.
  if (!b->host_envs || !b->host_level_off)
    return fail(FFSRL_E_ARG, "the plain refit needs host_envs and host_level_off");
  const ffsrl_env_desc* envs = b->host_envs;
  const int32_t* lvl = b->host_level_off;
  for (int i = eb; i < ee; ++i) {
    const ffsrl_env_desc& e = envs[i];
    for (int L = 0; L < e.nlvl; ++L) {
      int beg = lvl[e.lvlbase + L], end = lvl[e.lvlbase + L + 1];
      if (end <= beg) continue;
      hipLaunchKernelGGL(k_refit_level, dim3((end - beg + 255) / 256), dim3(256), 0, st, *b, e, beg, end);
    }
  }
  FFSRL_HIP(hipGetLastError());
  return 0;
}

int ffsrl_bvh_build(const float* verts, int32_t nv, const int32_t* faces, int32_t nf, int32_t bins, int32_t* left,
                    int32_t* first, int32_t* count, int32_t* level, int32_t* tri_map, int32_t* num_nodes,
                    int32_t* max_level) {
  if (!verts || !faces || !left || !first || !count || !level || !tri_map || !num_nodes || nv <= 0 || nf <= 0 ||
      bins < 2)
    return fail(FFSRL_E_ARG, "bvh_build: bad arguments");
  for (int64_t i = 0; i < 3 * (int64_t)nf; ++i)
    if (faces[i] < 0 || faces[i] >= nv) return fail(FFSRL_E_ARG, "bvh_build: face index %d out of range", faces[i]);
  Builder B;
  B.v = verts;
  B.f = faces;
  B.bins = bins;
  B.cen.resize(3 * (size_t)nf);
  for (int t = 0; t < nf; ++t)
    for (int k = 0; k < 3; ++k) {
      float s = B.vert(t, 0)[k] + B.vert(t, 1)[k];
      s = s + B.vert(t, 2)[k];
      B.cen[3 * t + k] = s / 3.0f;
    }
  B.tri.resize(nf);
  for (int t = 0; t < nf; ++t) B.tri[t] = t;
  B.box.resize(2 * (size_t)nf);
  B.left = left;
  B.first = first;
  B.count = count;
  B.level = level;
  for (int i = 0; i < 2 * nf; ++i) {
    left[i] = -1;
    first[i] = 0;
    count[i] = 0;
    level[i] = -1;
  }
  first[0] = 0;
  count[0] = nf;
  level[0] = 0;
  Box root;
  for (int i = 0; i < nv; ++i) root.grow(verts + 3 * (int64_t)i);
  B.box[0] = root;
  B.subdivide(0, 0);
  for (int t = 0; t < nf; ++t) tri_map[t] = B.tri[t];
  if (max_level) *max_level = B.max_level;
  *num_nodes = B.used;
  return 0;
}

}  // extern "C"
