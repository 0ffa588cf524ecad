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

namespace hcipy {
namespace {

// ---------------------------------------------------------------------------------------------
// Register-resident line FFT.  P threads share one line of N = P * E points; every thread holds
// E points in registers.  Stockham stages of radix E run as in-register DFTs with one LDS round
// trip between stages; N = E^a * RL, the last stage has radix RL < E (E / RL butterflies per
// thread) or is the a-th full stage when RL = 1.  On entry and on return slot q of v holds
// point tl + q * P (natural order), so a caller loads and stores coalesced rows and can chain
// transforms in registers.  sign = -1: forward (numpy fft), +1: inverse without 1/N.
// tw[t] = exp(-2 pi i t / N).  Every LDS write is bracketed by barriers, so the line may be
// reused between calls.  Point i of the line sits at LDS index i + i / 16 (one pad slot per 16
// points): stage 1 writes 16 consecutive points per thread, and without the pad the threads of a
// wave would hit the same banks; a line needs hr_lds_points(N) slots.

__host__ __device__ constexpr int hr_lds_points(int n) { return n + n / 16; }
__device__ inline int hr_pad(int i) { return i + (i >> 4); }

template <typename V, typename T>
__device__ inline V cmul_tw(V a, V w, T s) {
  w.y *= -s;  // tw holds exp(-i..): forward (s = -1) keeps it, inverse conjugates
  return cmul(a, w);
}

template <typename V, typename T>
__device__ inline void dft2(V& a, V& b) {
  const V t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <typename V, typename T>
__device__ inline void dft4(V& x0, V& x1, V& x2, V& x3, T s) {
  const V a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = cmul_si(csub(x1, x3), s);
  x0 = cadd(a, c);
  x2 = csub(a, c);
  x1 = cadd(b, d);
  x3 = csub(b, d);
}

// x[0..7] -> X[0..7] in natural order: two DFT4s and the radix-2 combine with w8^k.
template <typename V, typename T>
__device__ inline void dft8(V* x, T s) {
  dft4<V, T>(x[0], x[2], x[4], x[6], s);
  dft4<V, T>(x[1], x[3], x[5], x[7], s);
  const T h = (T)0.70710678118654752440;
  const V o1 = cmul(x[3], V{h, s * h});   // w8^1 O1
  const V o2 = cmul_si(x[5], s);          // w8^2 O2
  const V o3 = cmul(x[7], V{-h, s * h});  // w8^3 O3
  const V e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6], oo0 = x[1];
  x[0] = cadd(e0, oo0);
  x[4] = csub(e0, oo0);
  x[1] = cadd(e1, o1);
  x[5] = csub(e1, o1);
  x[2] = cadd(e2, o2);
  x[6] = csub(e2, o2);
  x[3] = cadd(e3, o3);
  x[7] = csub(e3, o3);
}

// x[0..15] -> X[0..15] in natural order: n = 4 n1 + n2, DFT4 over n1, twiddles w16^(n2 k1),
// DFT4 over n2, then X[k1 + 4 k2] is read from slot 4 k1 + k2.
template <typename V, typename T>
__device__ inline void dft16(V* x, T s) {
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<V, T>(x[n2], x[4 + n2], x[8 + n2], x[12 + n2], s);
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
  // slot 4 k1 + n2 *= w16^(n2 k1), w16^m = exp(s 2 pi i m / 16)
  x[5] = cmul(x[5], V{c1, s * s1});       // m = 1
  x[6] = cmul(x[6], V{h, s * h});         // m = 2
  x[7] = cmul(x[7], V{s1, s * c1});       // m = 3
  x[9] = cmul(x[9], V{h, s * h});         // m = 2
  x[10] = cmul_si(x[10], s);              // m = 4
  x[11] = cmul(x[11], V{-h, s * h});      // m = 6
  x[13] = cmul(x[13], V{s1, s * c1});     // m = 3
  x[14] = cmul(x[14], V{-h, s * h});      // m = 6
  x[15] = cmul(x[15], V{-c1, -s * s1});   // m = 9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<V, T>(x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3], s);
  V y[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) y[k1 + 4 * k2] = x[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) x[i] = y[i];
}

template <typename V, typename T, int R>
__device__ inline void dft_r(V* x, T s) {
  if constexpr (R == 2) {
    dft2<V, T>(x[0], x[1]);
  } else if constexpr (R == 4) {
    dft4<V, T>(x[0], x[1], x[2], x[3], s);
  } else if constexpr (R == 8) {
    dft8<V, T>(x, s);
  } else {
    static_assert(R == 16, "radix must be 2, 4, 8 or 16");
    dft16<V, T>(x, s);
  }
}

template <int N, int P>
struct HrShape {
  static constexpr int E = N / P;
  static constexpr int LOGN = __builtin_ctz(N);
  static constexpr int LOGE = __builtin_ctz(E);
  static constexpr int A = LOGN / LOGE;                // full stages of radix E
  static constexpr int RL = 1 << (LOGN - A * LOGE);    // radix of the last stage (1: none)
  static_assert(E == 2 || E == 4 || E == 8 || E == 16, "points per thread must be 2, 4, 8 or 16");
  static_assert(P * E == N, "N must be P * E");
};

// One LDS round trip of a line: every thread stores its E points at windex(r), and after a
// barrier loads slot r from hr_pad(tl + r P).  SPLIT moves the real parts and then the imaginary
// parts through a line of T (half the LDS, two more barriers); the line then holds
// hr_lds_points(N) values of T instead of complex values.
template <bool SPLIT, typename T, int E, int P, typename WI>
__device__ inline void hr_exchange(typename Cplx<T>::V (&v)[E], typename Cplx<T>::V* line, int tl, WI windex) {
  if constexpr (!SPLIT) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) line[windex(r)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = line[hr_pad(tl + r * P)];
  } else {
    T* l = reinterpret_cast<T*>(line);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) l[windex(r)] = v[r].x;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) v[r].x = l[hr_pad(tl + r * P)];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) l[windex(r)] = v[r].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) v[r].y = l[hr_pad(tl + r * P)];
  }
}

template <typename T, int N, int P, bool SPLIT = false>
__device__ void hr_fft(typename Cplx<T>::V (&v)[N / P], typename Cplx<T>::V* line, int tl,
                       const typename Cplx<T>::V* __restrict__ tw, int sign) {
  using V = typename Cplx<T>::V;
  using S = HrShape<N, P>;
  constexpr int E = S::E;
  constexpr int A = S::A;
  constexpr int RL = S::RL;
  const T s = (T)sign;
  // stage 1 (span 1): butterfly tl takes points tl + r P, no twiddles
  dft_r<V, T, E>(v, s);
  if constexpr (A == 1 && RL == 1) return;  // N == E: P == 1 and the outputs are in order
  hr_exchange<SPLIT, T, E, P>(v, line, tl, [&](int r) { return hr_pad(tl * E + r); });
  int Ns = E;
#pragma unroll
  for (int st = 1; st < A; ++st) {
    const int k = tl & (Ns - 1);
    const int step = N / (Ns * E);
#pragma unroll
    for (int r = 1; r < E; ++r) v[r] = cmul_tw(v[r], tw[(r * k * step) & (N - 1)], s);
    dft_r<V, T, E>(v, s);
    if (st == A - 1 && RL == 1) return;  // last stage: k == tl, outputs tl + r P are in order
    const int base = (tl - k) * E + k;
    hr_exchange<SPLIT, T, E, P>(v, line, tl, [&](int r) { return hr_pad(base + r * Ns); });
    Ns *= E;
  }
  if constexpr (RL > 1) {
    // span N / RL: butterfly j = tl + b P (b < E / RL) takes slots b + r (E / RL), twiddles w^(r j);
    // the last exchange left slot q = point tl + q P in v
    constexpr int NB = E / RL;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = tl + b * P;
      V x[RL];
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        x[r] = v[b + r * NB];
        if (r > 0) x[r] = cmul_tw(x[r], tw[(r * j) & (N - 1)], s);
      }
      dft_r<V, T, RL>(x, s);
#pragma unroll
      for (int r = 0; r < RL; ++r) v[b + r * NB] = x[r];
    }
  }
}

}  // namespace
}  // namespace hcipy
