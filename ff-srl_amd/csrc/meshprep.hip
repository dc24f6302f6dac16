// Host-side mesh preparation: greedy graph colouring of constraints.
//
// Two constraints conflict when they share a vertex.  Sorting each mesh's
// tets and edges by colour (stable within a colour) makes the 64 lanes of a
// wavefront touch 64 disjoint vertex sets, so the LDS reductions of the
// Jacobi deltas in the frame kernel never hit the same address twice within
// one instruction.  (The reference accumulates with global atomics in
// element order, integrator.py:105-109; the colouring only reorders the
// summation.)
#include <stdint.h>

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

}  // extern "C"
