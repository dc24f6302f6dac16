// Shared device helpers and the error channel of the C ABI.
//
// All HIP code is compiled with -ffp-contract=off and IEEE division / square
// root (no fast-math), so each helper rounds exactly as its C counterpart on
// the host.  The operation order of every helper is fixed and documented,
// because the BVH bounds and the ray-hit triangle ids must match a CPU
// evaluation bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/ffsrl.h"

namespace ffsrl {

// Thread-local message of the last failing call (ffsrl_last_error).
void set_error(const std::string& msg);
int fail(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

#define FFSRL_HIP(call)                                   \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return ::ffsrl::check_hip(e_, #call); \
  } while (0)

// FLOAT_EPSILON of the reference (sys.float_info.epsilon, a double epsilon
// used as a float threshold; integrator.py:8).
constexpr float kFeps = 2.220446049250313e-16f;

struct f3 {
  float x, y, z;
};

__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
// dot = (x*x' + y*y') + z*z'
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }
// normalize = component-wise division by the length when it is > 0, else 0.
__host__ __device__ inline f3 normalize(f3 a) {
  float l = length(a);
  if (l > 0.0f) return a / l;
  return f3{0.0f, 0.0f, 0.0f};
}
// min/max as compare-and-select with the first operand returned on "true";
// keeps the NaN and signed-zero behaviour identical on host and device.
__host__ __device__ inline float fmin_(float a, float b) { return a < b ? a : b; }
__host__ __device__ inline float fmax_(float a, float b) { return a > b ? a : b; }
__host__ __device__ inline f3 vmin(f3 a, f3 b) { return f3{fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
__host__ __device__ inline f3 vmax(f3 a, f3 b) { return f3{fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }

__device__ inline float4 ld4(const float* base, int64_t i) {
  return reinterpret_cast<const float4*>(base)[i];
}
__device__ inline void st4(float* base, int64_t i, float4 v) { reinterpret_cast<float4*>(base)[i] = v; }
__device__ inline int4 ldi4(const int32_t* base, int64_t i) { return reinterpret_cast<const int4*>(base)[i]; }
__device__ inline int2 ldi2(const int32_t* base, int64_t i) { return reinterpret_cast<const int2*>(base)[i]; }

// Loads of state that other workgroups on the same XCD wrote during the
// kernel: nontemporal loads skip the CU's L1, which is not coherent between
// CUs, and read the XCD's L2, where stores land once `s_waitcnt vmcnt(0)`
// has returned.  (An agent-scope L2 invalidate would also make them visible,
// at the cost of the L2's contents.)
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ inline float4 ld4c(const float* base, int64_t i) {
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(base) + i);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline float ldfc(const float* base, int64_t i) { return __builtin_nontemporal_load(base + i); }
__device__ inline unsigned long long ldu64c(const unsigned long long* base, int64_t i) {
  return __builtin_nontemporal_load(base + i);
}
// COH: state written by other workgroups of a shared env (ld4c), else ld4.
template <bool COH>
__device__ inline float4 ldp(const float* P, int64_t i) {
  return COH ? ld4c(P, i) : ld4(P, i);
}
// Waits until this wavefront's vector memory operations (loads, stores and
// atomics) have completed: s_waitcnt vmcnt(0).
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }
// Hardware id of the XCD (XCC) the calling wavefront runs on.
__device__ inline int xcc_id() { return (int)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20); }

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace ffsrl
