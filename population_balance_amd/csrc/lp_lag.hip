// lp_lag.hip -- the Lagrangian scale-dependent model's 42 test filters fused with the
// products that feed them and the contractions that consume them
// (lagrange_Sdep.f90:129-434 with interpolag_Sdep.f90:62-365).
//
//   k_lag_r2c    forms field q of a row from u, v, w, S_ij, |S| in registers and does the
//                x transform (four-step FFT + real split), writing only the kx columns
//                below the filters' cutoff (the 21 inputs are never stored as volumes);
//   col_op       (lp_fft.hip) filters all 21 fields along y at both widths in one pass;
//   row_c2r      (lp_fft.hip) transforms them back along x: 42 filtered volumes;
//   k_lag_relax  interpolates the old averages at the backtraced point, contracts the
//                42 filtered values and relaxes the averages, one thread per point.
#include "lp_fft4.hpp"
#include "lp_lag.hpp"

namespace lp {

template <typename T> struct LagIn {
  const T *u, *v, *w;
  const T *S[6];
  const T *magS;
};
struct F4c {
  const double *p[4];
};
struct F4 {
  double *p[4];
};

// lagrange_Sdep.f90:129-155,192-242: values of input field q at the two points i, i+1
// (i even).  Order: 0-2 u, v, w on w-nodes, 3-8 the products 11,12,13,22,23,33,
// 9-14 S_ij, 15-20 |S| S_ij.  q is uniform over the wavefront.
template <typename T>
__device__ __forceinline__ typename V2<T>::t lag_in_pair(int q, const LagIn<T> &in, const Geo &g, long i, int k) {
  using v2 = typename V2<T>::t;
  auto ld2 = [](const T *f, long o) { return *reinterpret_cast<const v2 *>(f + o); };
  auto hv = [&](const T *f) -> v2 {  // u, v averaged to the w-node (jz = 1 stays)
    v2 a = ld2(f, i);
    if (k == 0) return a;
    v2 b = ld2(f, i - g.np);
    return v2{T(0.5) * (a.x + b.x), T(0.5) * (a.y + b.y)};
  };
  auto wn = [&]() -> v2 {  // w; at jz = 1 a quarter of w(2)
    if (k == 0) {
      v2 a = ld2(in.w, i + g.np);
      return v2{T(0.25) * a.x, T(0.25) * a.y};
    }
    return ld2(in.w, i);
  };
  auto mul = [](v2 a, v2 b) { return v2{a.x * b.x, a.y * b.y}; };
  switch (q) {
    case 0: return hv(in.u);
    case 1: return hv(in.v);
    case 2: return wn();
    case 3: { v2 a = hv(in.u); return mul(a, a); }
    case 4: return mul(hv(in.u), hv(in.v));
    case 5: return mul(hv(in.u), wn());
    case 6: { v2 a = hv(in.v); return mul(a, a); }
    case 7: return mul(hv(in.v), wn());
    case 8: { v2 a = wn(); return mul(a, a); }
    default:
      if (q < 15) return ld2(in.S[q - 9], i);
      return mul(ld2(in.magS, i), ld2(in.S[q - 15], i));
  }
}

// One wavefront transforms one field for 64/P rows; the four wavefronts of a block take
// four fields at a time, six rounds cover the 21.
template <typename T, int M>
__global__ __launch_bounds__(256) void k_lag_r2c(LagIn<T> in, Geo g, cx<T> *__restrict__ spec, int nc, long pitch,
                                                 const cx<T> *__restrict__ twM, const cx<T> *__restrict__ twN) {
  constexpr int P = row_lanes(M);
  using F = Fft4<T, M, P, -1>;
  constexpr int Q = F::Q;
  constexpr int RB = 64 / P;
  constexpr int SS = slot_stride(F::SCR > M ? F::SCR : M);
  __shared__ cx<T> lds[4 * RB * SS];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l / P, t = l % P;
  const long nrows = (long)g.ny * g.nz;
  const long row = (long)blockIdx.x * RB + r;
  const bool valid = row < nrows;
  const int k = valid ? (int)(row / g.ny) : 0, y = valid ? (int)(row % g.ny) : 0;
  const long base = (long)k * g.np + (long)y * g.ld;
  cx<T> *sc = lds + (wv * RB + r) * SS;
  for (int it = 0; it < 6; ++it) {
    const int q = __builtin_amdgcn_readfirstlane(4 * it + wv);
    const bool act = valid && q < 21;
    cx<T> z[Q];
#pragma unroll
    for (int m = 0; m < Q; ++m) {
      typename V2<T>::t a{T(0), T(0)};
      if (act) a = lag_in_pair<T>(q, in, g, base + 2 * (t + P * m), k);
      z[m] = mk<T>(a.x, a.y);
    }
    F::stage1(z, t, sc, twM);
    slot_sync<P>();
    cx<T> a[F::G][P];
    bool ok[F::G];
#pragma unroll
    for (int j = 0; j < F::G; ++j) ok[j] = F::stage2(j, t, sc, a[j]);
    slot_sync<P>();
#pragma unroll
    for (int j = 0; j < F::G; ++j)
      if (ok[j]) {
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) sc[t + P * j + Q * k1] = a[j][k1];
      }
    slot_sync<P>();
    if (act) {  // split (as k_row_r2c), leading nc columns
      cx<T> *dst = spec + (((long)q * g.nz + k) * g.ny + y) * pitch;
      for (int kk = t; kk < nc; kk += P) {
        cx<T> za = sc[kk % M], zb = conjg(sc[(M - kk) % M]);
        cx<T> e = za + zb, o = za - zb;
        cx<T> wo = cmul(twN[kk % (2 * M)], o);
        dst[kk] = mk<T>(T(0.5) * (e.x + wo.y), T(0.5) * (e.y - wo.x));
      }
    }
    slot_sync<P>();
  }
}

// lagrange_Sdep.f90:257-434 with interpolag_Sdep.f90:62-365, one thread per point:
// bar / hat hold the 21 fields filtered at 2 and 4 delta as volumes (field q at q*nv);
// the old averages are gathered at the backtraced point, relaxed and written to Fnew.
template <typename T>
__global__ __launch_bounds__(256) void k_lag_relax(const T *__restrict__ bar, const T *__restrict__ hat, F4c Fold,
                                                   F4 Fnew, const T *__restrict__ ul, const T *__restrict__ vl,
                                                   const T *__restrict__ wl, T *__restrict__ Cs,
                                                   T *__restrict__ Beta, Geo g, LagParams P, LagInterp I) {
  const long nv = g.np * g.nz;
  const long i = gtid();
  if (i >= nv) return;
  const int k = (int)(i / g.np);
  const int r = (int)(i % g.np);
  const int y = r / g.ld, x = r % g.ld;
  const double *const Fo[4] = {Fold.p[0], Fold.p[1], Fold.p[2], Fold.p[3]};
  double fi[4];
  if (x < g.nx) {
    interp_point<T>(Fo, fi, ul, vl, wl, g, i, x, y, k, I.cs, I.dt, I.dx, I.dy, I.dz);
  } else {  // interpolag_Sdep.f90 leaves the padding columns alone
#pragma unroll
    for (int q = 0; q < 4; ++q) fi[q] = Fo[q][i];
  }
  double lm, mm, qn, nn;
  contract_point<T>([&](int q) -> T { return bar[q * nv + i]; }, P.cst, P.tf1_2, lm, mm);
  contract_point<T>([&](int q) -> T { return hat[q * nv + i]; }, P.cst, P.tf2_2, qn, nn);
  T cs, b;
  relax_point<T>(lm, mm, qn, nn, fi, x >= g.ld - 2, k == g.nz - 1, P, cs, b);
#pragma unroll
  for (int q = 0; q < 4; ++q) Fnew.p[q][i] = fi[q];
  Cs[i] = cs;
  Beta[i] = b;
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
#define LP_SMALL_ROWS(X) X(16) X(32) X(48) X(64) X(96) X(128)

template <typename T, int M>
static hipError_t launch_lag_r2c(const LagInputs &in, Geo g, void *spec, int nc, long pitch, hipStream_t s,
                                 Dt dt) {
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN || nc < 1 || nc > M + 1 || pitch < nc) return hipErrorInvalidValue;
  LagIn<T> a;
  a.u = (const T *)in.u; a.v = (const T *)in.v; a.w = (const T *)in.w;
  for (int q = 0; q < 6; ++q) a.S[q] = (const T *)in.S[q];
  a.magS = (const T *)in.magS;
  constexpr int RB = 64 / row_lanes(M);
  const long rows = (long)g.ny * g.nz;
  hipLaunchKernelGGL((k_lag_r2c<T, M>), dim3((unsigned)((rows + RB - 1) / RB)), dim3(256), 0, s, a, g,
                     (cx<T> *)spec, nc, pitch, (const cx<T> *)twM, (const cx<T> *)twN);
  return hipGetLastError();
}

hipError_t lag_r2c(Dt dt, const LagInputs &in, Geo g, void *spec, int nc, long pitch, hipStream_t s) {
#define CASE(m)                                                                          \
  case m:                                                                                \
    return dt == F64 ? launch_lag_r2c<double, m>(in, g, spec, nc, pitch, s, dt)          \
                     : launch_lag_r2c<float, m>(in, g, spec, nc, pitch, s, dt);
  switch (g.nx / 2) {
    LP_SMALL_ROWS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

template <typename T>
static hipError_t launch_lag_relax(const void *bar, const void *hat, const double *const Fold[4],
                                   double *const Fnew[4], const void *ul, const void *vl, const void *wl,
                                   void *Cs, void *Beta, Geo g, const LagParams &P, const LagInterp &I,
                                   hipStream_t s) {
  F4c fo;
  F4 fn;
  for (int q = 0; q < 4; ++q) { fo.p[q] = Fold[q]; fn.p[q] = Fnew[q]; }
  const long n = g.np * g.nz;
  hipLaunchKernelGGL((k_lag_relax<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const T *)bar,
                     (const T *)hat, fo, fn, (const T *)ul, (const T *)vl, (const T *)wl, (T *)Cs, (T *)Beta, g,
                     P, I);
  return hipGetLastError();
}

hipError_t lag_relax(Dt dt, const void *bar, const void *hat, const double *const Fold[4], double *const Fnew[4],
                     const void *ul, const void *vl, const void *wl, void *Cs, void *Beta, Geo g,
                     const LagParams &P, const LagInterp &I, hipStream_t s) {
  return dt == F64 ? launch_lag_relax<double>(bar, hat, Fold, Fnew, ul, vl, wl, Cs, Beta, g, P, I, s)
                   : launch_lag_relax<float>(bar, hat, Fold, Fnew, ul, vl, wl, Cs, Beta, g, P, I, s);
}

}  // namespace lp
