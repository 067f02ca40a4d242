// lp_device.hpp -- device-side building blocks for gfx950 (CDNA4, wave64): complex
// arithmetic and the small in-register DFTs (radix 2, 3, 4) that the four-step FFT of
// lp_fft4.hpp is built from.
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

}  // namespace lp
