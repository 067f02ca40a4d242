// lp_fft4.hpp -- register-resident FFT building blocks for gfx950 (wave64).
//
// Dft<T,R,SIGN> is a fully unrolled mixed-radix DFT of R values held in registers
// (R | 144: 2, 3, 4, 6, 8, 9, 12, 16, 18, 24, 48), with constant twiddles folded at
// compile time from kW144.
//
// Fft4<T,N,P,SIGN> is the four-step ("six-step" without the transposes in memory)
// factorisation N = P*Q over P cooperating lanes:
//   lane t holds x[t + P*q], q < Q                       (stage-1 input)
//   stage 1: Q-point DFT in registers, twiddle W_N^(t*k2), write to LDS (k2, n1 = t)
//   barrier
//   stage 2: for k2 = t + P*j: P-point DFT over n1 of the LDS row k2, giving
//            X[k2 + Q*k1], k1 < P
// so every transform costs one LDS round trip.  W = exp(SIGN*2*pi*i/N); SIGN = -1 is
// the forward (FFTW_FORWARD) transform, +1 the unnormalised backward one.
#pragma once
#include "lp_device.hpp"

namespace lp {

// exp(-2*pi*i*m/144), rounded to binary64 from 40-digit values
__device__ constexpr double kW144[144][2] = {
#include "lp_w144.inc"
};

template <typename T> struct V2;
template <> struct V2<double> { using t = double2; };
template <> struct V2<float> { using t = float2; };

// lanes per row transform of M complex points (P) -- M = P*Q
__host__ __device__ constexpr int row_lanes(int M) {
  return M == 16 ? 4 : M == 24 ? 4 : M == 32 ? 4 : M == 48 ? 4 : M == 64 ? 8 : M == 72 ? 6
       : M == 96 ? 8 : M == 128 ? 8 : M == 144 ? 12 : M == 192 ? 8 : 0;
}

// LDS stride (elements) of a transform slot holding n elements: rounded so that
// consecutive slots start 4 elements apart modulo 32
__host__ __device__ constexpr int slot_stride(int n) { return ((n + 27) / 32) * 32 + 4; }

// Barrier for the LDS slot of one transform done by P adjacent lanes: when P divides the
// wavefront the slot belongs to one wavefront and a wavefront-scope barrier orders its
// LDS accesses; otherwise the whole block synchronises.
template <int P>
__device__ __forceinline__ void slot_sync() {
  if constexpr (64 % P == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <typename T, int SIGN>
__device__ __forceinline__ cx<T> tw_const(int m, int R) {
  const int idx = (m % R) * (144 / R);
  const T c = (T)kW144[idx][0];
  const T s = (T)(SIGN < 0 ? kW144[idx][1] : -kW144[idx][1]);
  return mk<T>(c, s);
}

// a * exp(SIGN*2*pi*i*m/R) with the trivial roots done without multiplications
template <typename T, int SIGN>
__device__ __forceinline__ cx<T> rot(cx<T> a, int m, int R) {
  m %= R;
  if (m == 0) return a;
  if (4 * m == R) return SIGN < 0 ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x);      // -+i
  if (2 * m == R) return mk<T>(-a.x, -a.y);
  if (4 * m == 3 * R) return SIGN < 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);  // +-i
  return cmul(a, tw_const<T, SIGN>(m, R));
}

template <typename T, int R, int SIGN> struct Dft {
  static __device__ __forceinline__ void run(cx<T> (&v)[R]) {
    if constexpr (R == 1) {
      return;
    } else if constexpr (R == 2 || R == 3 || R == 4) {
      SmallDft<T, R, SIGN>::run(v);
    } else {
      constexpr int A = (R % 4 == 0) ? 4 : (R % 2 == 0) ? 2 : 3;
      constexpr int B = R / A;
      cx<T> y[A][B];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) {
        cx<T> s[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) s[n2] = v[n1 + A * n2];
        Dft<T, B, SIGN>::run(s);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) y[n1][k2] = rot<T, SIGN>(s[k2], n1 * k2, R);
      }
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) {
        cx<T> s[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) s[n1] = y[n1][k2];
        Dft<T, A, SIGN>::run(s);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = s[k1];
      }
    }
  }
};

template <typename T, int N, int P, int SIGN> struct Fft4 {
  static constexpr int Q = N / P;
  static constexpr int G = (Q + P - 1) / P;  // stage-2 groups per lane
  static constexpr int LD = P + 1;            // padded LDS row (one k2)
  static constexpr int SCR = Q * LD;          // LDS elements per transform
  static_assert(P * Q == N, "N must be P*Q");

  // v: x[t + P*q] on entry; consumed.  twN: exp(-2*pi*i*j/N), j < N.
  static __device__ __forceinline__ void stage1(cx<T> (&v)[Q], int t, cx<T> *sc, const cx<T> *__restrict__ twN) {
    Dft<T, Q, SIGN>::run(v);
#pragma unroll
    for (int k2 = 1; k2 < Q; ++k2) {
      cx<T> w = twN[t * k2];
      if (SIGN > 0) w.y = -w.y;
      v[k2] = cmul(v[k2], w);
    }
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) sc[k2 * LD + t] = v[k2];
  }

  // after a barrier: a[k1] = X[k2 + Q*k1] for k2 = t + P*j; false when k2 >= Q
  static __device__ __forceinline__ bool stage2(int j, int t, const cx<T> *sc, cx<T> (&a)[P]) {
    const int k2 = t + P * j;
    if (k2 >= Q) return false;
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) a[n1] = sc[k2 * LD + n1];
    Dft<T, P, SIGN>::run(a);
    return true;
  }
};

}  // namespace lp
