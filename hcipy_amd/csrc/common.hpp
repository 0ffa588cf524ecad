// Shared definitions for the gfx950 kernels of the Fourier-propagation path.
//
// Everything here is CDNA4 code: 64-lane wavefronts, MFMA intrinsics, LDS tiles.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef double doublex4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));  // 16-byte register vector

namespace hcipy {

// Error plumbing: every C-ABI entry point returns one of these codes and keeps a
// thread-local message for hcipy_get_last_error().
enum Status : int {
  kOk = 0,
  kInvalidValue = 1,
  kHipError = 2,
  kOutOfRange = 3,
  kUnsupported = 4,
};

void set_last_error(const char* fmt, ...);

// Returns kHipError (and records the message) if `e` is not hipSuccess.
int check_hip(hipError_t e, const char* what);

// Plan creation allocates device memory and waits on the stream, which a graph capture cannot
// record: returns kInvalidValue with a message naming `what` when `stream` is capturing.
int refuse_under_capture(hipStream_t stream, const char* what);

#define HCIPY_HIP(expr)                                        \
  do {                                                         \
    int _st = ::hcipy::check_hip((expr), #expr);               \
    if (_st != ::hcipy::kOk) return _st;                       \
  } while (0)

#define HCIPY_REQUIRE(cond, ...)                               \
  do {                                                         \
    if (!(cond)) {                                             \
      ::hcipy::set_last_error(__VA_ARGS__);                    \
      return ::hcipy::kInvalidValue;                           \
    }                                                          \
  } while (0)

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }
inline int64_t ceil_div(int64_t v, int64_t m) { return (v + m - 1) / m; }

// float -> half with the native conversion (round to nearest even).  The empty asm makes the
// float opaque, so that a preceding double -> float rounding is not merged into one double ->
// half conversion, which the compiler would emulate in ~25 integer instructions.
__device__ inline _Float16 f32_to_f16(float f) {
  __asm__ volatile("" : "+v"(f));
  return (_Float16)f;
}

// float -> bfloat16, round to nearest even (v_cvt_pk_bf16_f32), with the float made opaque as in
// f32_to_f16 so that a preceding double -> float rounding stays a separate step.
__device__ inline __bf16 f32_to_bf16(float f) {
  __asm__ volatile("" : "+v"(f));
  return (__bf16)f;
}

// Wave-level reductions over the 64 lanes of a CDNA wavefront.
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace hcipy
