// lp_device.hpp -- device-side building blocks for gfx950 (CDNA4, wave64):
// complex arithmetic and an LDS-resident Stockham FFT with compile-time radix plan.
//
// The FFT is the auto-sort Stockham formulation of Govindaraju et al., "High
// Performance Discrete Fourier Transforms on Graphics Processors" (SC'08): each stage
// reads R elements strided by N/R, applies twiddles exp(sign*2*pi*i*(j mod Ns)*r/(Ns*R))
// and an R-point DFT, and writes them at (j/Ns)*Ns*R + (j mod Ns) + r*Ns.  Radices are
// 4, then 2, then 3 (every length used here is 2^a 3^b).  Twiddles come from a table
// tw[j] = exp(-2*pi*i*j/N) computed on the host in binary64.
#pragma once

// the reference's BOGUS marker (param.f90) for levels a routine leaves undefined
#define LP_BOGUS (-1234567890.0)
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lp {

template <typename T> struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T> __device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ cx<T> conjg(cx<T> a) { return mk(a.x, -a.y); }
template <typename T> __device__ __forceinline__ cx<T> scal(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
// multiply by i*s
template <typename T> __device__ __forceinline__ cx<T> mul_i(cx<T> a, T s) { return mk(-a.y * s, a.x * s); }

// ---- small DFTs (in registers) ----
template <typename T, int R, int SIGN> struct SmallDft;

template <typename T, int SIGN> struct SmallDft<T, 2, SIGN> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <typename T, int SIGN> struct SmallDft<T, 4, SIGN> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    // X_q = sum_r v_r w^(rq), w = exp(SIGN*2*pi*i/4) = SIGN*i
    cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
    cx<T> b0 = v[1] + v[3], b1 = v[1] - v[3];
    // (SIGN*i) * b1
    cx<T> jb1 = SIGN < 0 ? mk(b1.y, -b1.x) : mk(-b1.y, b1.x);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + jb1;
    v[3] = a1 - jb1;
  }
};

template <typename T, int SIGN> struct SmallDft<T, 3, SIGN> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    // w = exp(SIGN*2*pi*i/3) = -1/2 + SIGN*i*sqrt(3)/2
    const T c1 = T(-0.5);
    const T s1 = T(SIGN) * T(0.86602540378443864676372317075294);
    cx<T> t1 = v[1] + v[2];
    cx<T> t2 = v[1] - v[2];
    cx<T> m = mk(v[0].x + c1 * t1.x, v[0].y + c1 * t1.y);
    // i*s1*t2
    cx<T> j = mk(-s1 * t2.y, s1 * t2.x);
    v[0] = v[0] + t1;
    v[1] = m + j;
    v[2] = m - j;
  }
};

__host__ __device__ constexpr int next_radix(int rem) {
  return (rem % 4 == 0) ? 4 : (rem % 2 == 0) ? 2 : (rem % 3 == 0) ? 3 : rem;
}

// One Stockham stage over a transform of length N living in src (LDS), written to dst.
// nt threads cooperate, this thread is t.
template <typename T, int N, int Ns, int R, int SIGN>
__device__ __forceinline__ void stockham_stage(const cx<T> *__restrict__ src, cx<T> *__restrict__ dst,
                                               int t, int nt, const cx<T> *__restrict__ tw) {
  constexpr int NB = N / R;
  for (int j = t; j < NB; j += nt) {
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[j + r * NB];
    const int kk = j % Ns;
    if constexpr (Ns > 1) {
      constexpr int step = N / (Ns * R);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = tw[kk * r * step];
        if (SIGN > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    SmallDft<T, R, SIGN>::run(v);
    const int base = (j / Ns) * Ns * R + kk;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
  }
}

// Full transform: data starts in a; buffers a and b ping-pong; returns the buffer
// holding the result.  Every thread of the block must call it (it synchronises).
template <typename T, int N, int Ns, int SIGN>
__device__ __forceinline__ cx<T> *stockham_run(cx<T> *a, cx<T> *b, int t, int nt,
                                               const cx<T> *__restrict__ tw) {
  if constexpr (Ns >= N) {
    return a;
  } else {
    constexpr int R = next_radix(N / Ns);
    stockham_stage<T, N, Ns, R, SIGN>(a, b, t, nt, tw);
    __syncthreads();
    return stockham_run<T, N, Ns * R, SIGN>(b, a, t, nt, tw);
  }
}

template <typename T, int N, int SIGN>
__device__ __forceinline__ cx<T> *fft_lds(cx<T> *a, cx<T> *b, int t, int nt, const cx<T> *__restrict__ tw) {
  return stockham_run<T, N, 1, SIGN>(a, b, t, nt, tw);
}

}  // namespace lp
