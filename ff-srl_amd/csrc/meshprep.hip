// Host-side mesh preparation: the order in which the frame kernel visits a
// mesh's constraints.
//
// The kernel's Jacobi sums are integers, so the order changes neither the
// result nor its reproducibility (the reference accumulates with global
// atomics in element order, integrator.py:105-109); it only changes speed:
//   * spatial locality: the position gathers of a wavefront should hit few
//     cache lines, so constraints are sorted along a Morton curve of their
//     centroids;
//   * LDS address conflicts: two lanes of one ds_add to the same vertex are
//     serialised, so within each chunk of consecutive constraints the order is
//     a stable sort by a greedy colour (two constraints conflict when they
//     share a vertex).
// ffsrl_color_constraints colours a whole list; ffsrl_order_constraints
// builds the frame kernel's order (Morton, then colour within chunks).
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"

using namespace ffsrl;

extern "C" {

int ffsrl_color_constraints(const int32_t* elems, int64_t n, int32_t arity, int32_t nv, int32_t* color,
                            int32_t* num_colors) {
  constexpr int kWords = 4;  // up to 256 colours
  if (!elems || !color || !num_colors || n < 0 || arity < 1 || arity > 8 || nv < 0)
    return fail(FFSRL_E_ARG, "color_constraints: bad arguments");
  std::vector<uint64_t> used((size_t)nv * kWords, 0);
  int32_t ncol = 0;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t m[kWords] = {0, 0, 0, 0};
    for (int k = 0; k < arity; ++k) {
      const int32_t v = elems[i * arity + k];
      if (v < 0 || v >= nv) return fail(FFSRL_E_ARG, "color_constraints: vertex %d out of range", v);
      for (int w = 0; w < kWords; ++w) m[w] |= used[(size_t)v * kWords + w];
    }
    int c = -1;
    for (int w = 0; w < kWords && c < 0; ++w)
      if (~m[w]) c = w * 64 + __builtin_ctzll(~m[w]);
    if (c < 0) return fail(FFSRL_E_LIMIT, "color_constraints: more than %d colours needed", 64 * kWords);
    color[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
    for (int k = 0; k < arity; ++k) used[(size_t)elems[i * arity + k] * kWords + c / 64] |= 1ull << (c % 64);
  }
  *num_colors = ncol;
  return 0;
}

static uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

int ffsrl_order_constraints(const int32_t* elems, int64_t n, int32_t arity, const float* verts, int32_t nv,
                            int32_t chunk, int64_t* perm) {
  if (!elems || !verts || !perm || n < 0 || arity < 1 || arity > 8 || nv < 0 || chunk < 1)
    return fail(FFSRL_E_ARG, "order_constraints: bad arguments");
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int32_t v = 0; v < nv; ++v)
    for (int a = 0; a < 3; ++a) {
      const float x = verts[3 * v + a];
      if (v == 0 || x < lo[a]) lo[a] = x;
      if (v == 0 || x > hi[a]) hi[a] = x;
    }
  std::vector<uint32_t> key((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t k = 0;
    for (int a = 0; a < 3; ++a) {
      double c = 0.0;
      for (int j = 0; j < arity; ++j) {
        const int32_t v = elems[i * arity + j];
        if (v < 0 || v >= nv) return fail(FFSRL_E_ARG, "order_constraints: vertex %d out of range", v);
        c += verts[3 * v + a];
      }
      c /= arity;
      const double ext = (double)hi[a] - lo[a];
      const double t = ext > 0 ? (c - lo[a]) / ext : 0.0;
      k |= spread10((uint32_t)std::min(1023.0, std::max(0.0, t * 1023.0))) << a;
    }
    key[(size_t)i] = k;
  }
  std::vector<int64_t> order((size_t)n);
  std::iota(order.begin(), order.end(), int64_t{0});
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[(size_t)a] < key[(size_t)b]; });
  // greedy colour within each chunk, then a stable sort of the chunk by colour
  constexpr int kWords = 4;
  std::vector<uint64_t> used((size_t)nv * kWords, 0);
  std::vector<int32_t> col((size_t)std::min<int64_t>(n, chunk));
  std::vector<int64_t> idx(col.size());
  for (int64_t a = 0; a < n; a += chunk) {
    const int64_t m = std::min<int64_t>(chunk, n - a);
    for (int64_t i = 0; i < m; ++i) {
      const int32_t* e = elems + order[(size_t)(a + i)] * arity;
      uint64_t mask[kWords] = {0, 0, 0, 0};
      for (int j = 0; j < arity; ++j)
        for (int w = 0; w < kWords; ++w) mask[w] |= used[(size_t)e[j] * kWords + w];
      int c = -1;
      for (int w = 0; w < kWords && c < 0; ++w)
        if (~mask[w]) c = w * 64 + __builtin_ctzll(~mask[w]);
      if (c < 0) c = 64 * kWords;  // more colours than tracked: keep the Morton position
      col[(size_t)i] = c;
      if (c < 64 * kWords)
        for (int j = 0; j < arity; ++j) used[(size_t)e[j] * kWords + c / 64] |= 1ull << (c % 64);
    }
    for (int64_t i = 0; i < m; ++i) {
      const int32_t* e = elems + order[(size_t)(a + i)] * arity;
      for (int j = 0; j < arity; ++j)
        for (int w = 0; w < kWords; ++w) used[(size_t)e[j] * kWords + w] = 0;
      idx[(size_t)i] = i;
    }
    std::stable_sort(idx.begin(), idx.begin() + m, [&](int64_t x, int64_t y) { return col[(size_t)x] < col[(size_t)y]; });
    for (int64_t i = 0; i < m; ++i) perm[a + i] = order[(size_t)(a + idx[(size_t)i])];
  }
  return 0;
}

// Partition of one mesh for the partitioned frame: part p takes tets
// [p nt / P, (p+1) nt / P) and edges [p ne / P, (p+1) ne / P) of the current
// order; its vertices are those its constraints in use touch, plus the
// untouched vertices, dealt round-robin in increasing id, and, with the
// collision pass, the vertices of the surface triangles it is given: a
// triangle goes to the lowest part holding its first vertex.  A part's
// vertices are numbered locally in increasing mesh id.  A vertex is flagged
// SHARED when several parts hold it and OWNER on the lowest part that holds
// it.  Three sweeps over the constraints with a per-vertex stamp (membership
// counts, then triangle extras, then the lists): O(4 nt + 2 ne + 3 nf + nv)
// plus the sort of each part's vertex list.
int ffsrl_partition_mesh(const int32_t* tets, int64_t nt, const int32_t* edges, int64_t ne, const int32_t* tris,
                         int64_t nf, int32_t nv, int32_t parts, int32_t use_tets, int32_t use_edges,
                         int32_t use_tris, int32_t* rows, int32_t* verts, int64_t verts_cap, int64_t* n_verts,
                         uint64_t* tets16, uint32_t* edges16, uint64_t* tris16, int32_t* max_local) {
  if (parts < 1 || nv < 0 || nt < 0 || ne < 0 || nf < 0 || !rows || !verts || !n_verts || !max_local ||
      (use_tets && nt && (!tets || !tets16)) || (use_edges && ne && (!edges || !edges16)) ||
      (use_tris && nf && (!tris || !tris16)))
    return fail(FFSRL_E_ARG, "partition_mesh: bad arguments");
  const int64_t P = parts;
  auto tbeg = [&](int64_t p) { return p * nt / P; };
  auto ebeg = [&](int64_t p) { return p * ne / P; };
  auto in_range = [&](int32_t v) { return v >= 0 && v < nv; };
  if (use_tets)
    for (int64_t i = 0; i < 4 * nt; ++i)
      if (!in_range(tets[i])) return fail(FFSRL_E_ARG, "partition_mesh: a tet vertex id outside [0, %d)", nv);
  if (use_edges)
    for (int64_t i = 0; i < 2 * ne; ++i)
      if (!in_range(edges[i])) return fail(FFSRL_E_ARG, "partition_mesh: an edge vertex id outside [0, %d)", nv);
  if (use_tris)
    for (int64_t f = 0; f < nf; ++f)
      for (int k = 0; k < 3; ++k)
        if (!in_range(tris[4 * f + k]))
          return fail(FFSRL_E_ARG, "partition_mesh: a triangle vertex id outside [0, %d)", nv);
  std::vector<int32_t> stamp((size_t)nv, -1), holders((size_t)nv, 0), first((size_t)nv, -1);
  // visits the vertices of part p's constraints in use (a vertex once per constraint)
  auto each = [&](int64_t p, auto&& fn) {
    if (use_tets)
      for (int64_t t = tbeg(p); t < tbeg(p + 1); ++t)
        for (int k = 0; k < 4; ++k) fn(tets[4 * t + k]);
    if (use_edges)
      for (int64_t e = ebeg(p); e < ebeg(p + 1); ++e)
        for (int k = 0; k < 2; ++k) fn(edges[2 * e + k]);
  };
  // memberships from the constraints, then the untouched vertices
  for (int64_t p = 0; p < P; ++p)
    each(p, [&](int32_t v) {
      if (stamp[(size_t)v] != (int32_t)p) {
        stamp[(size_t)v] = (int32_t)p;
        if (holders[(size_t)v]++ == 0) first[(size_t)v] = (int32_t)p;
      }
    });
  std::vector<std::vector<int32_t>> loose((size_t)P);
  int64_t nl = 0;
  for (int32_t v = 0; v < nv; ++v)
    if (holders[(size_t)v] == 0) {
      const int32_t p = (int32_t)(nl++ % P);
      loose[(size_t)p].push_back(v);
      holders[(size_t)v] = 1;
      first[(size_t)v] = p;
    }
  // triangles by part (the lowest part holding the first vertex), in leaf order within a part
  std::vector<int64_t> tri_off((size_t)P + 1, 0), tri_list;
  if (use_tris) {
    for (int64_t f = 0; f < nf; ++f) tri_off[(size_t)first[(size_t)tris[4 * f]] + 1] += 1;
    for (int64_t p = 0; p < P; ++p) tri_off[(size_t)p + 1] += tri_off[(size_t)p];
    tri_list.resize((size_t)nf);
    std::vector<int64_t> fill(tri_off.begin(), tri_off.end() - 1);
    for (int64_t f = 0; f < nf; ++f) tri_list[(size_t)fill[(size_t)first[(size_t)tris[4 * f]]]++] = f;
    // the triangles' vertices join their part
    std::fill(stamp.begin(), stamp.end(), -1);
    for (int64_t p = 0; p < P; ++p) {
      each(p, [&](int32_t v) { stamp[(size_t)v] = (int32_t)p; });
      for (int32_t v : loose[(size_t)p]) stamp[(size_t)v] = (int32_t)p;
      for (int64_t j = tri_off[(size_t)p]; j < tri_off[(size_t)p + 1]; ++j)
        for (int k = 0; k < 3; ++k) {
          const int32_t v = tris[4 * tri_list[(size_t)j] + k];
          if (stamp[(size_t)v] != (int32_t)p) {
            stamp[(size_t)v] = (int32_t)p;
            holders[(size_t)v] += 1;
            first[(size_t)v] = std::min(first[(size_t)v], (int32_t)p);
          }
        }
    }
  }
  std::fill(stamp.begin(), stamp.end(), -1);
  std::vector<int32_t> local((size_t)nv, 0), list;
  int64_t off = 0;
  int32_t most = 0;
  for (int64_t p = 0; p < P; ++p) {
    list.clear();
    auto add = [&](int32_t v) {
      if (stamp[(size_t)v] != (int32_t)p) {
        stamp[(size_t)v] = (int32_t)p;
        list.push_back(v);
      }
    };
    each(p, add);
    for (int32_t v : loose[(size_t)p]) add(v);
    if (use_tris)
      for (int64_t j = tri_off[(size_t)p]; j < tri_off[(size_t)p + 1]; ++j)
        for (int k = 0; k < 3; ++k) add(tris[4 * tri_list[(size_t)j] + k]);
    std::sort(list.begin(), list.end());
    const int64_t n = (int64_t)list.size();
    if (off + n > verts_cap) return fail(FFSRL_E_ARG, "partition_mesh: verts holds %lld entries", (long long)verts_cap);
    for (int64_t i = 0; i < n; ++i) {
      const int32_t v = list[(size_t)i];
      local[(size_t)v] = (int32_t)i;
      uint32_t f = (uint32_t)v;
      if (holders[(size_t)v] > 1) f |= FFSRL_PART_SHARED;
      if (first[(size_t)v] == (int32_t)p) f |= FFSRL_PART_OWNER;
      verts[off + i] = (int32_t)f;
    }
    if (use_tets)
      for (int64_t t = tbeg(p); t < tbeg(p + 1); ++t) {
        uint64_t w = 0;
        for (int k = 0; k < 4; ++k) w |= (uint64_t)(uint16_t)local[(size_t)tets[4 * t + k]] << (16 * k);
        tets16[t] = w;
      }
    if (use_edges)
      for (int64_t e = ebeg(p); e < ebeg(p + 1); ++e)
        edges16[e] = (uint32_t)(uint16_t)local[(size_t)edges[2 * e]] |
                     ((uint32_t)(uint16_t)local[(size_t)edges[2 * e + 1]] << 16);
    if (use_tris)
      for (int64_t j = tri_off[(size_t)p]; j < tri_off[(size_t)p + 1]; ++j) {
        uint64_t w = 0;
        for (int k = 0; k < 3; ++k) w |= (uint64_t)(uint16_t)local[(size_t)tris[4 * tri_list[(size_t)j] + k]] << (16 * k);
        tris16[j] = w;
      }
    int32_t* r = rows + 8 * p;
    r[0] = (int32_t)tbeg(p);
    r[1] = (int32_t)tbeg(p + 1);
    r[2] = (int32_t)ebeg(p);
    r[3] = (int32_t)ebeg(p + 1);
    r[4] = (int32_t)off;
    r[5] = (int32_t)n;
    r[6] = use_tris ? (int32_t)tri_off[(size_t)p] : 0;
    r[7] = use_tris ? (int32_t)(tri_off[(size_t)p + 1] - tri_off[(size_t)p]) : 0;
    off += n;
    most = std::max(most, (int32_t)n);
  }
  if (!use_tets && tets16)
    for (int64_t t = 0; t < nt; ++t) tets16[t] = 0;
  if (!use_edges && edges16)
    for (int64_t e = 0; e < ne; ++e) edges16[e] = 0;
  if (!use_tris && tris16)
    for (int64_t f = 0; f < nf; ++f) tris16[f] = 0;
  *n_verts = off;
  *max_local = most;
  return 0;
}

}  // extern "C"
