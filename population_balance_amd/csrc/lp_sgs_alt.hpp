// The other dynamic SGS models of sgs_stag.f90:188-198 on the generic spectral path:
// model 2 std_dynamic.f90 and model 3 scaledep_dynamic.f90 (one Cs^2 per level from
// plane sums), model 4 lagrange_Ssim.f90 with interpolag_Ssim.f90 (Lagrangian averages
// of L.M and M.M only).
//
// The filter inputs are formed as volumes (k_dyn_inputs), test-filtered level by level
// in the context's spectral passes, and contracted point by point.  Plane sums run in one
// workgroup per level with a fixed lane order and LDS tree.  Field order of the input
// and filtered volumes: 0-2 u, v, w on the w-node, 3-8 the products 11 12 13 22 23 33,
// 9-14 S_ij, 15-20 |S| S_ij.  Model 3 also filters with the 4-delta filter: u, v, w, the
// products of the 2-delta filtered velocities (its Q_ij, scaledep_dynamic.f90:100-120),
// S_ij and |S| S_ij.
#pragma once
#include "lp_kernels.hpp"
#include "lp_scalar.hpp"

namespace lp {

constexpr int kDynIn = 21;

struct DynVol {  // 21 volumes of the field type
  const void *p[kDynIn];
};

// std_dynamic.f90:36-62 + :80-100: the inputs that are not S_ij itself (fields 0-8 and
// 15-20; 9-14 are the strain volumes).  out[q]: volume q of 21 (9-14 unused).  Binary32:
// u, v on the w-node enter less the level's first value (L_ij and Q_ij are unchanged by a
// constant: the 2-delta and 4-delta filters pass kx = ky = 0 with gain 1 and the 4-delta
// passband lies inside the 2-delta one), binary64 keeps the reference's values.
template <typename T>
__global__ void k_dyn_inputs(const double *__restrict__ u, const double *__restrict__ v,
                             const double *__restrict__ w, const T *__restrict__ S0, const T *__restrict__ S1,
                             const T *__restrict__ S2, const T *__restrict__ S3, const T *__restrict__ S4,
                             const T *__restrict__ S5, T *__restrict__ out, Geo g, int shift) {
  const long nv = g.np * g.nz;
  const long i = gtid();
  if (i >= nv) return;
  const int k = (int)(i / g.np);
  const long r = i - (long)k * g.np;
  auto hv = [&](const double *f, long j) { return k == 0 ? f[j] : 0.5 * (f[j] + f[j - g.np]); };
  double cu = 0.0, cv = 0.0;
  if (shift) {
    cu = hv(u, (long)k * g.np);
    cv = hv(v, (long)k * g.np);
  }
  const T ub = (T)(hv(u, i) - cu), vb = (T)(hv(v, i) - cv);
  const T wb = (T)(k == 0 ? 0.25 * w[g.np + r] : w[i]);
  out[0 * nv + i] = ub; out[1 * nv + i] = vb; out[2 * nv + i] = wb;
  out[3 * nv + i] = ub * ub; out[4 * nv + i] = ub * vb; out[5 * nv + i] = ub * wb;
  out[6 * nv + i] = vb * vb; out[7 * nv + i] = vb * wb; out[8 * nv + i] = wb * wb;
  const T s[6] = {S0[i], S1[i], S2[i], S3[i], S4[i], S5[i]};
  const T sm = smag(s[0], s[1], s[2], s[3], s[4], s[5]);
#pragma unroll
  for (int q = 0; q < 6; ++q) out[(15 + q) * nv + i] = sm * s[q];
}

// scaledep_dynamic.f90:100-105: products of the 2-delta filtered velocities
template <typename T>
__global__ void k_dyn_qprod(const T *__restrict__ ub, const T *__restrict__ vb, const T *__restrict__ wb,
                            T *__restrict__ out, long nv) {
  const long i = gtid();
  if (i >= nv) return;
  const T a = ub[i], b = vb[i], c = wb[i];
  out[0 * nv + i] = a * a; out[1 * nv + i] = a * b; out[2 * nv + i] = a * c;
  out[3 * nv + i] = b * b; out[4 * nv + i] = b * c; out[5 * nv + i] = c * c;
}

template <typename T>
__device__ __forceinline__ T dget(const DynVol &V, int q, long i) {
  return static_cast<const T *>(V.p[q])[i];
}

// a . b in the reference's order: 11 + 22 + 33 + 2 (12 + 13 + 23)
__device__ __forceinline__ double dot6(const double (&a)[6], const double (&b)[6]) {
  return a[0] * b[0] + a[3] * b[3] + a[5] * b[5] + 2.0 * (a[1] * b[1] + a[2] * b[2] + a[4] * b[4]);
}

// the filtered fields of one width at point i: L_ij (from fields 0-8; the 4-delta Q_ij
// from the filtered products of the 2-delta velocities), S_ij bar, |S| S_ij bar, |S|bar
template <typename T>
__device__ __forceinline__ void dyn_point(const DynVol &F, long i, double (&L)[6], double (&Sb)[6],
                                          double (&SSb)[6], double &sbar) {
  const double ub = dget<T>(F, 0, i), vb = dget<T>(F, 1, i), wb = dget<T>(F, 2, i);
  L[0] = dget<T>(F, 3, i) - ub * ub; L[1] = dget<T>(F, 4, i) - ub * vb; L[2] = dget<T>(F, 5, i) - ub * wb;
  L[3] = dget<T>(F, 6, i) - vb * vb; L[4] = dget<T>(F, 7, i) - vb * wb; L[5] = dget<T>(F, 8, i) - wb * wb;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    Sb[q] = dget<T>(F, 9 + q, i);
    SSb[q] = dget<T>(F, 15 + q, i);
  }
  sbar = sqrt(2.0 * (Sb[0] * Sb[0] + Sb[3] * Sb[3] + Sb[5] * Sb[5] + 2.0 * (Sb[1] * Sb[1] + Sb[2] * Sb[2] + Sb[4] * Sb[4])));
}

// scaledep_dynamic.f90:156-198: the ten plane sums of level blockIdx.x (sum[k][10])
template <typename T>
__global__ __launch_bounds__(1024) void k_dyn_sums(DynVol B, DynVol H, Geo g, double *__restrict__ sums) {
  __shared__ double red[1024];
  const int k = blockIdx.x;
  const long nxy = (long)g.nx * g.ny;
  double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long q = threadIdx.x; q < nxy; q += blockDim.x) {
    const long i = (long)k * g.np + (q / g.nx) * g.ld + q % g.nx;
    double L[6], Sb[6], SSb[6], sb;
    dyn_point<T>(B, i, L, Sb, SSb, sb);
    s[0] += sb * dot6(Sb, L);
    s[1] += dot6(SSb, L);
    s[2] += dot6(SSb, SSb);
    s[3] += 0.5 * (sb * sb) * (sb * sb);
    s[4] += sb * dot6(Sb, SSb);
    dyn_point<T>(H, i, L, Sb, SSb, sb);
    s[5] += sb * dot6(Sb, L);
    s[6] += dot6(SSb, L);
    s[7] += dot6(SSb, SSb);
    s[8] += 0.5 * (sb * sb) * (sb * sb);
    s[9] += sb * dot6(Sb, SSb);
  }
#pragma unroll
  for (int m = 0; m < 10; ++m) {
    const double t = block_sum1024(s[m], red);
    if (threadIdx.x == 0) sums[k * 10 + m] = t;
  }
}

// :199-209 and sgsmodule.f90:47-69: the quintic of the sums and its Newton root (from 7.5,
// a step below 0.001 ends it, 100 steps without that give 1); one lane per level
static __global__ void k_dyn_beta(const double *__restrict__ sums, double *__restrict__ beta, int nz, double delta,
                           double nxy) {
  const int k = (int)gtid();
  if (k >= nz) return;
  const double *s = sums + k * 10;
  const double d2 = delta * delta, c2 = (2.0 * d2) * (2.0 * d2);
  const double a1 = -2.0 * d2 * 4.0 * s[0] / nxy, b1 = -2.0 * d2 * s[1] / nxy;
  const double c1 = c2 * s[2] / nxy, dd1 = c2 * 16.0 * s[3] / nxy, e1 = 2.0 * c2 * 4.0 * s[4] / nxy;
  const double a2 = -2.0 * d2 * 16.0 * s[5] / nxy, b2 = -2.0 * d2 * s[6] / nxy;
  const double cc2 = c2 * s[7] / nxy, dd2 = c2 * 256.0 * s[8] / nxy, e2 = 2.0 * c2 * 16.0 * s[9] / nxy;
  const double A0 = b2 * c1 - b1 * cc2, A1 = a1 * cc2 - b2 * e1, A2 = b2 * dd1 + b1 * e2 - a2 * c1;
  const double A3 = a2 * e1 - a1 * e2, A4 = -a2 * dd1 - b1 * dd2, A5 = a1 * dd2;
  double r = 0.5 * (0.0 + 15.0), out = 1.0;
  for (int j = 0; j < 100; ++j) {
    const double f = A0 + r * (A1 + r * (A2 + r * (A3 + r * (A4 + r * A5))));
    const double df = A1 + r * (2.0 * A2 + r * (3.0 * A3 + r * (4.0 * A4 + r * (5.0 * A5))));
    const double dx = f / df;
    r = r - dx;
    if (fabs(dx) < 0.001) {
      out = r;
      break;
    }
  }
  beta[k] = out;
}

// std_dynamic.f90:115-132 / scaledep_dynamic.f90:211-224: Cs^2 of level blockIdx.x from the
// plane sums of L.M and M.M with M_ij = 2 delta^2 (|S|S_ij bar - 4 beta |S|bar S_ij bar),
// clipped at zero, over the whole plane (sgs_stag.f90:192, 196); with beta (model 3) the
// level's beta is left in Beta (scaledep_dynamic keeps it local; nothing else reads it)
template <typename T>
__global__ __launch_bounds__(1024) void k_dyn_cs(DynVol B, const double *__restrict__ beta, Geo g, double cst,
                                                 T *__restrict__ Cs, T *__restrict__ Beta) {
  __shared__ double red[1024];
  __shared__ double val;
  const int k = blockIdx.x;
  const long nxy = (long)g.nx * g.ny;
  const double b = beta ? beta[k] : 1.0, fourb = 4.0 * b;
  double slm = 0.0, smm = 0.0;
  for (long q = threadIdx.x; q < nxy; q += blockDim.x) {
    const long i = (long)k * g.np + (q / g.nx) * g.ld + q % g.nx;
    double L[6], Sb[6], SSb[6], sb, M[6];
    dyn_point<T>(B, i, L, Sb, SSb, sb);
#pragma unroll
    for (int m = 0; m < 6; ++m) M[m] = cst * (SSb[m] - fourb * sb * Sb[m]);
    slm += dot6(L, M);
    smm += dot6(M, M);
  }
  slm = block_sum1024(slm, red);
  smm = block_sum1024(smm, red);
  if (threadIdx.x == 0) val = fmax(0.0, slm / smm);
  __syncthreads();
  for (long i = threadIdx.x; i < g.np; i += blockDim.x) {
    Cs[(long)k * g.np + i] = (T)val;
    if (beta) Beta[(long)k * g.np + i] = (T)b;
  }
}

// lagrange_Ssim.f90:178-192: L.M and M.M (M with 4 |S|bar S_ij bar) at every point; zero in
// the padding columns
template <typename T>
__global__ void k_dyn_contract(DynVol B, Geo g, double cst, T *__restrict__ lm, T *__restrict__ mm) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int x = (int)(i % g.ld);
  double a = 0.0, c = 0.0;
  if (x < g.nx) {
    double L[6], Sb[6], SSb[6], sb, M[6];
    dyn_point<T>(B, i, L, Sb, SSb, sb);
#pragma unroll
    for (int m = 0; m < 6; ++m) M[m] = cst * (SSb[m] - 4.0 * sb * Sb[m]);
    a = dot6(L, M);
    c = dot6(M, M);
  }
  lm[i] = (T)a;
  mm[i] = (T)c;
}

// interpolag_Ssim.f90:118-156: F_LM, F_MM carried back along the trajectory (one lane per
// node; columns >= nx keep their old values).  The lag velocities as interpolag_Sdep's
// (u, v on the w-node, w at jz = 1 a quarter of w(2)); at the wall only zp(2) += min(zp(2), 0)
// and zp(1) += max(zp(1), 0) (:180-185), no clamp.
template <typename T>
__global__ void k_interpolag_ssim(const double *__restrict__ F0, const double *__restrict__ F1,
                                  double *__restrict__ I0, double *__restrict__ I1, const T *__restrict__ ul,
                                  const T *__restrict__ vl, const T *__restrict__ wl, Geo g, double cs, double sx,
                                  double sy, double sz) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int k = (int)(i / g.np);
  const long r = i - (long)k * g.np;
  const int y = (int)(r / g.ld), x = (int)(r % g.ld);
  if (x >= g.nx) {
    I0[i] = F0[i];
    I1[i] = F1[i];
    return;
  }
  const double ut = (double)ul[i] / cs, vt = (double)vl[i] / cs;
  double ulag = ut, vlag = vt;
  if (k > 0) {
    ulag = 0.5 * (ut + (double)ul[i - g.np] / cs);
    vlag = 0.5 * (vt + (double)vl[i - g.np] / cs);
  }
  const double xp = -ulag * sx, yp = -vlag * sy;
  double zp;
  if (k == 0) {
    const double z1 = -(0.25 * ((double)wl[g.np + r] / cs)) * sz;
    zp = z1 + fmax(z1, 0.0);
  } else if (k == 1) {
    const double z2 = -((double)wl[i] / cs) * sz;
    zp = z2 + fmin(z2, 0.0);
  } else {
    zp = -((double)wl[i] / cs) * sz;
  }
  const int addx = signbit(xp) ? -1 : 1, addy = signbit(yp) ? -1 : 1, addz = signbit(zp) ? -1 : 1;
  const double cxp = fabs(xp), cyp = fabs(yp), czp = fabs(zp);
  const double fx = 1.0 - cxp, fy = 1.0 - cyp, fz = 1.0 - czp;
  int xa = (x + addx + g.nx) % g.nx, ya = (y + addy + g.ny) % g.ny, ka = k + addz;
  ka = ka < 0 ? 0 : (ka > g.nz - 1 ? g.nz - 1 : ka);
  const long o00 = (long)k * g.np + (long)y * g.ld, o0a = (long)ka * g.np + (long)y * g.ld;
  const long oa0 = (long)k * g.np + (long)ya * g.ld, oaa = (long)ka * g.np + (long)ya * g.ld;
  const double *const F[2] = {F0, F1};
  double *const O[2] = {I0, I1};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double *f = F[q];
    O[q][i] = fx * fy * (f[o00 + x] * fz + f[o0a + x] * czp) + fx * cyp * (f[oa0 + x] * fz + f[oaa + x] * czp) +
              cxp * fy * (f[o00 + xa] * fz + f[o0a + xa] * czp) + cxp * cyp * (f[oa0 + xa] * fz + f[oaa + xa] * czp);
  }
}

// lagrange_Ssim.f90:193-218: the first call's initial averages (init: F_MM = M.M,
// F_LM = 0.025 M.M, the last two columns 1), the relaxation, Cs^2 = F_LM / F_MM (zero in the
// last two columns), floors 1e-32; REAL(4) conversions where the reference makes them
// (casts).  Beta = 1.
template <typename T>
__global__ void k_ssim_relax(const T *__restrict__ lm_, const T *__restrict__ mm_, double *__restrict__ FLM,
                             double *__restrict__ FMM, T *__restrict__ Cs, T *__restrict__ Beta, Geo g, int init,
                             int casts, double opftdelta, double lagran_dt) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int x = (int)(i % g.ld);
  const bool edge = x >= g.ld - 2;
  const double lm = lm_[i], mm = mm_[i];
  const double tiny = casts ? (double)1e-32f : 1e-32;
  double flm = FLM[i], fmm = FMM[i];
  if (init) {
    fmm = mm;
    flm = 0.025 * mm;
    if (edge) { fmm = 1.0; flm = 1.0; }
  }
  double Tn = casts ? (double)fmaxf((float)(flm * fmm), 1e-32f) : fmax(flm * fmm, tiny);
  Tn = opftdelta * pow(Tn, -1.0 / 8.0);
  const double dum = lagran_dt / Tn, epsi = dum / (1.0 + dum);
  flm = (epsi * lm + (1.0 - epsi) * flm);
  fmm = (epsi * mm + (1.0 - epsi) * fmm);
  flm = casts ? (double)fmaxf(1e-32f, (float)flm) : fmax(tiny, flm);
  double cs = flm / fmm;
  if (edge) cs = 0.0;
  cs = casts ? (double)fmaxf(1e-32f, (float)cs) : fmax(tiny, cs);
  FLM[i] = flm;
  FMM[i] = fmm;
  Cs[i] = (T)cs;
  Beta[i] = T(1);
}

}  // namespace lp
