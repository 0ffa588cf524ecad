// Complex 2-vectors and the in-LDS power-of-two line FFT shared by the FourierFilter passes and
// the chirp-z (zoom FFT) passes.  Radix-4 Stockham passes (one radix-2 pass when log2 N is odd)
// in LDS, in place, with all loads of a pass in registers before its stores; twiddles come from an
// FP64-generated table exp(-2 pi i t / N).
#pragma once

#include <hip/hip_runtime.h>

namespace hcipy {
namespace {

// Complex values as native 2-vectors (register-resident; the HIP float2/double2 structs are
// copied through memory by the compiler).
typedef float cf32_t __attribute__((ext_vector_type(2)));
typedef double cf64_t __attribute__((ext_vector_type(2)));
template <typename T> struct Cplx;
template <> struct Cplx<float> { using V = cf32_t; };
template <> struct Cplx<double> { using V = cf64_t; };

template <typename V>
__device__ inline V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename V>
__device__ inline V cadd(V a, V b) {
  V r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename V>
__device__ inline V csub(V a, V b) {
  V r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
// a * (s i), s = +-1
template <typename V, typename T>
__device__ inline V cmul_si(V a, T s) {
  V r;
  r.x = -s * a.y;
  r.y = s * a.x;
  return r;
}

// In-place FFT of one LDS line of length N (power of two) by TPL cooperating threads.
// sign = -1: forward (numpy fft), +1: inverse without 1/N.  tw[t] = exp(-2 pi i t / N).
template <typename T, int N, int TPL>
__device__ void line_fft(typename Cplx<T>::V* line, int tl, const typename Cplx<T>::V* __restrict__ tw, int sign,
                         int sync_barrier_all) {
  using V = typename Cplx<T>::V;
  constexpr int LOG2N = __builtin_ctz(N);
  constexpr int NR4 = LOG2N / 2;            // radix-4 passes
  constexpr bool R2 = (LOG2N % 2) != 0;     // one final radix-2 pass
  constexpr int B4 = (N / 4 + TPL - 1) / TPL;  // radix-4 butterflies per thread
  constexpr int B2 = (N / 2 + TPL - 1) / TPL;
  const T s = (T)sign;
  int Ns = 1;   // span of the sub-transforms done so far (a power of 4)
  int lg = 0;   // log2(Ns): index arithmetic by shifts and masks
  // FP32 lines: passes unrolled (index arithmetic folds to constants; measured faster); FP64
  // lines keep the loop (unrolling raises register pressure and was measured slower)
  constexpr int UNR = sizeof(T) == 4 ? NR4 : 1;
#pragma unroll UNR
  for (int p = 0; p < NR4; ++p) {
    V v[B4][4];
#pragma unroll
    for (int b = 0; b < B4; ++b) {
      const int j = tl + b * TPL;
      if (j < N / 4) {
        const int k = j & (Ns - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[b][r] = line[j + r * (N / 4)];
        if (Ns > 1) {
          const int step = N >> (lg + 2);
#pragma unroll
          for (int r = 1; r < 4; ++r) {
            V w = tw[(r * k * step) & (N - 1)];
            w.y *= -s;  // tw holds exp(-i..); forward (s=-1) keeps it, inverse conjugates
            v[b][r] = cmul(v[b][r], w);
          }
        }
        const V a = cadd(v[b][0], v[b][2]);
        const V bb = csub(v[b][0], v[b][2]);
        const V c = cadd(v[b][1], v[b][3]);
        const V d = cmul_si(csub(v[b][1], v[b][3]), s);
        v[b][0] = cadd(a, c);
        v[b][2] = csub(a, c);
        v[b][1] = cadd(bb, d);
        v[b][3] = csub(bb, d);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B4; ++b) {
      const int j = tl + b * TPL;
      if (j < N / 4) {
        const int k = j & (Ns - 1);
        const int o = ((j >> lg) << (lg + 2)) + k;
#pragma unroll
        for (int r = 0; r < 4; ++r) line[o + r * Ns] = v[b][r];
      }
    }
    __syncthreads();
    Ns *= 4;
    lg += 2;
  }
  if constexpr (R2) {
    V v[B2][2];
#pragma unroll
    for (int b = 0; b < B2; ++b) {
      const int j = tl + b * TPL;
      if (j < N / 2) {
        const int k = j & (Ns - 1);
        v[b][0] = line[j];
        v[b][1] = line[j + N / 2];
        if (Ns > 1) {
          V w = tw[(k * (N >> (lg + 1))) & (N - 1)];
          w.y *= -s;
          v[b][1] = cmul(v[b][1], w);
        }
        const V a = v[b][0];
        v[b][0] = cadd(a, v[b][1]);
        v[b][1] = csub(a, v[b][1]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B2; ++b) {
      const int j = tl + b * TPL;
      if (j < N / 2) {
        const int k = j & (Ns - 1);
        const int o = ((j >> lg) << (lg + 1)) + k;
        line[o] = v[b][0];
        line[o + Ns] = v[b][1];
      }
    }
    __syncthreads();
  }
  (void)sync_barrier_all;
}

}  // namespace
}  // namespace hcipy
