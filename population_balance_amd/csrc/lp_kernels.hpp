// lp_kernels.hpp -- pointwise / stencil / gather / tridiagonal kernels of the LES step
// (everything that is not an FFT pass).  Each kernel cites the reference lines it
// implements.  T is the field type (double for LP_FP64, float for LP_MIXED); the
// Lagrangian F_* state, the pressure solve, p and the velocities u, v, w are always double.
#pragma once
#include "lp_device.hpp"

namespace lp {

struct Geo {
  int nx, ny, nz, ld;
  long np;  // ld*ny
};

__device__ __forceinline__ long gtid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }

// ddz_uv.f90:39-57 ; level 1 untouched.  TI: type of f (binary64 velocities), the
// difference is formed in TI and stored in T.
template <typename T, typename TI = T>
__global__ void k_ddz_uv(T *__restrict__ dfdz, const TI *__restrict__ f, Geo g, TI cst) {
  long i = gtid();
  if (i >= g.np * g.nz) return;
  int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (k == g.nz - 1) dfdz[i] = T(0);
  else if (k >= 1 && x < g.nx) dfdz[i] = (T)(cst * (f[i] - f[i - g.np]));
}

// ddz_w.f90:24-44
template <typename T, typename TI = T>
__global__ void k_ddz_w(T *__restrict__ dfdz, const TI *__restrict__ f, Geo g, TI cst) {
  long i = gtid();
  if (i >= g.np * g.nz) return;
  int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (k == g.nz - 1) dfdz[i] = T(0);
  else if (x < g.nx) dfdz[i] = (T)(cst * (f[i + g.np] - f[i]));
}

// the kx = 0 column of x-spectra in physical y ([nplanes][ny][pitch] complex): add a
// per-plane value to the real part of every y, i.e. to the plane's mean (the inverse of
// row_r2c's shift on a spectrum that keeps the row means there)
template <typename T>
__global__ void k_add_dc(cx<T> *__restrict__ spec, long pitch, int ny, int nplanes, const double *__restrict__ dc) {
  long i = gtid();
  if (i >= (long)ny * nplanes) return;
  const int k = (int)(i / ny);
  cx<T> &e = spec[i * pitch];
  e.x = (T)((double)e.x + dc[k]);
}

// filt_da.f90's zeroing of the kx = nx/2 and ky = ny/2 modes ("oddballs") on a binary64
// field, exactly, without a round trip through binary32 transforms.  Per level those
// modes sum to
//   (-1)^x A(y) + (-1)^y B(x) - (-1)^(x+y) C,
//   A(y) = sum_x (-1)^x f(x, y) / nx,  B(x) = sum_y (-1)^y f(x, y) / ny,
//   C = sum_y (-1)^y A(y) / ny  (the corner mode, counted in both sums).
constexpr int kNyqMax = 384;  // largest nx, ny of the transforms

// The sums in two passes that read f once: k_nyq_part gives each of kNyqSlices workgroups
// of a level a contiguous slice of its rows (4 waves, each taking every 4th row of the
// slice, kNyqRU rows at a time with their loads issued together), which yields A(y) of its
// rows (a wave reduction) and its partial alternating column sums P[k][s][x] (the 4 waves'
// partials added in LDS); k_nyq_fin adds the slices' partials in slice order for B and
// forms C from A.  Every sum has a fixed order: the result does not depend on timing.
constexpr int kNyqSlices = 16, kNyqRU = 4;
static __global__ __launch_bounds__(256) void k_nyq_part(const double *__restrict__ f, double *__restrict__ A,
                                                         double *__restrict__ P, Geo g) {
  __shared__ double colp[4][kNyqMax];
  const int lane = threadIdx.x % 64, grp = threadIdx.x / 64;
  const int k = blockIdx.x / kNyqSlices, sl = blockIdx.x % kNyqSlices;
  const int y0 = sl * g.ny / kNyqSlices, y1 = (sl + 1) * g.ny / kNyqSlices;
  const double *p = f + (long)k * g.np;
  constexpr int NQ = kNyqMax / 64;
  double col[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) col[q] = 0.0;
  for (int yb = y0 + grp; yb < y1; yb += 4 * kNyqRU) {
    double v[kNyqRU][NQ];
#pragma unroll
    for (int u = 0; u < kNyqRU; ++u) {
      const int y = yb + 4 * u;
      const double *r = p + (long)(y < y1 ? y : y0) * g.ld;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int x = lane + 64 * q;
        v[u][q] = (y < y1 && x < g.nx) ? r[x] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kNyqRU; ++u) {
      const int y = yb + 4 * u;
      double ra = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        ra += (lane & 1) ? -v[u][q] : v[u][q];  // x and lane have the same parity
        col[q] += (y & 1) ? -v[u][q] : v[u][q];
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ra += __shfl_xor(ra, o, 64);
      if (lane == 0 && y < y1) A[(long)k * g.ny + y] = ra / g.nx;
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int x = lane + 64 * q;
    if (x < g.nx) colp[grp][x] = col[q];
  }
  __syncthreads();
  for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < 4; ++r) s += colp[r][x];
    P[((long)k * kNyqSlices + sl) * g.nx + x] = s;
  }
}
static __global__ __launch_bounds__(256) void k_nyq_fin(const double *__restrict__ A, const double *__restrict__ P,
                                                        double *__restrict__ B, double *__restrict__ Cc, Geo g) {
  const int k = blockIdx.x;
  for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
    double s = 0.0;
    for (int sl = 0; sl < kNyqSlices; ++sl) s += P[((long)k * kNyqSlices + sl) * g.nx + x];
    B[(long)k * g.nx + x] = s / g.ny;
  }
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    double c = 0.0;
    for (int y = lane; y < g.ny; y += 64) {
      const double a = A[(long)k * g.ny + y];
      c += (y & 1) ? -a : a;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) Cc[k] = c / g.ny;
  }
}

// f -= the oddball modes (padding columns x >= nx set to zero, as a back-transform leaves them)
static __global__ void k_nyq_remove(double *__restrict__ f, const double *__restrict__ A, const double *__restrict__ B,
                             const double *__restrict__ Cc, Geo g, int nlev) {
  const long i = gtid();
  if (i >= g.np * nlev) return;
  const int x = (int)(i % g.ld), y = (int)((i / g.ld) % g.ny), k = (int)(i / g.np);
  if (x >= g.nx) {
    f[i] = 0.0;
    return;
  }
  const double sx = (x & 1) ? -1.0 : 1.0, sy = (y & 1) ? -1.0 : 1.0;
  f[i] = f[i] - (sx * A[(long)k * g.ny + y] + sy * B[(long)k * g.nx + x] - sx * sy * Cc[k]);
}

// wallstress.f90:101-185 ('wall', uniform z0, d0): denom = log(zl / z0) less psi_m of
// obukhov (psi: the psi_m plane, ob[2]: phi_m; both null with theta off: 0 and 1)
template <typename T>
__global__ void k_wallstress(const T *__restrict__ u1, const T *__restrict__ v1, const double *__restrict__ u,
                             const double *__restrict__ v, T *__restrict__ txz, T *__restrict__ tyz,
                             T *__restrict__ dudz, T *__restrict__ dvdz, T *__restrict__ ust, Geo g,
                             double denom, double zl, double vonk, const double *__restrict__ psi,
                             const double *__restrict__ ob) {
  long i = gtid();
  if (i >= g.np) return;
  int x = (int)(i % g.ld);
  if (x >= g.nx) return;
  const double phi_m = ob ? ob[2] : 1.0;
  double a = u1[i], b = v1[i];
  double u_avg = sqrt(a * a + b * b);
  double us = u_avg * vonk / (psi ? denom - psi[i] : denom);
  double cst = -(us * us) / u_avg;
  txz[i] = (T)(cst * a);
  tyz[i] = (T)(cst * b);
  double uu = u[i], vv = v[i];
  double du = us / (zl * vonk) * uu / u_avg * phi_m;
  double dv = us / (zl * vonk) * vv / u_avg * phi_m;
  dudz[i] = (T)(uu == 0.0 ? 0.0 : du);
  dvdz[i] = (T)(vv == 0.0 ? 0.0 : dv);
  ust[i] = (T)us;
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, long n, T v) {
  long i = gtid();
  if (i < n) a[i] = v;
}

template <typename T> __device__ __forceinline__ T smag(T s11, T s12, T s13, T s22, T s23, T s33) {
  return sqrt(T(2) * (s11 * s11 + s22 * s22 + s33 * s33 + T(2) * (s12 * s12 + s13 * s13 + s23 * s23)));
}

// sgs_stag.f90:49-125 strain rates: uvp nodes at jz = 1, w-nodes above
template <typename T>
__global__ void k_strain(const T *__restrict__ dudx, const T *__restrict__ dudy, const T *__restrict__ dudz,
                         const T *__restrict__ dvdx, const T *__restrict__ dvdy, const T *__restrict__ dvdz,
                         const T *__restrict__ dwdx, const T *__restrict__ dwdy, const T *__restrict__ dwdz,
                         T *__restrict__ S11, T *__restrict__ S12, T *__restrict__ S13, T *__restrict__ S22,
                         T *__restrict__ S23, T *__restrict__ S33, T *__restrict__ Smag, Geo g) {
  long i = gtid();
  if (i >= g.np * g.nz) return;
  int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (x >= g.nx) return;
  T ux, uy, uz, vx, vy, vz, wx, wy, wz;
  const T h = T(0.5);
  if (k == 0) {
    ux = dudx[i]; uy = dudy[i]; uz = dudz[i];
    vx = dvdx[i]; vy = dvdy[i]; vz = dvdz[i];
    wx = h * (dwdx[i] + dwdx[i + g.np]);
    wy = h * (dwdy[i] + dwdy[i + g.np]);
    wz = dwdz[i];
  } else {
    long m = i - g.np;
    ux = h * (dudx[i] + dudx[m]); uy = h * (dudy[i] + dudy[m]); uz = dudz[i];
    vx = h * (dvdx[i] + dvdx[m]); vy = h * (dvdy[i] + dvdy[m]); vz = dvdz[i];
    wx = dwdx[i]; wy = dwdy[i]; wz = h * (dwdz[i] + dwdz[m]);
  }
  const T s11 = ux, s12 = h * (uy + vx), s13 = h * (uz + wx), s22 = vy, s23 = h * (vz + wy), s33 = wz;
  S11[i] = s11; S12[i] = s12; S13[i] = s13;
  S22[i] = s22; S23[i] = s23; S33[i] = s33;
  if (Smag) Smag[i] = smag(s11, s12, s13, s22, s23, s33);  // sgs_stag.f90:215 |S|
}

// sgs_stag.f90:130-134  u_lag = u_lag + u (all columns)
template <typename T>
__global__ void k_lag_acc(T *__restrict__ ul, T *__restrict__ vl, T *__restrict__ wl, const double *__restrict__ u,
                          const double *__restrict__ v, const double *__restrict__ w, long n) {
  long i = gtid();
  if (i >= n) return;
  ul[i] = (T)((double)ul[i] + u[i]);
  vl[i] = (T)((double)vl[i] + v[i]);
  wl[i] = (T)((double)wl[i] + w[i]);
}


// sgs_stag.f90:137-177 (mixing length) and :215-227 (|S|, Nu_t, dissip, magS) of level k
template <typename T>
__device__ __forceinline__ T nut_point(const T *__restrict__ S11, const T *__restrict__ S12,
                                       const T *__restrict__ S13, const T *__restrict__ S22,
                                       const T *__restrict__ S23, const T *__restrict__ S33,
                                       const T *__restrict__ Cs, long i, int k, int model, double delta, double Co,
                                       int nnn, double vonk, double dz, T &s, T &c, T &l2) {
  double l;
  if (model == 1) {
    double zz = (k == 0) ? 0.5 * dz : (double)k * dz;
    double n = (double)nnn;
    l = pow(pow(Co, n) * pow(vonk * zz, -n) + pow(delta, -n), -1.0 / n);
  } else {
    l = delta;
  }
  l2 = (T)(l * l);
  s = smag(S11[i], S12[i], S13[i], S22[i], S23[i], S33[i]);
  c = Cs[i];
  return s * c * l2;
}

// sgs_stag.f90:137-177, 215-227 (Nu_t, dissip, magS) and :254-325, 391-394 (the six
// stresses) in one pass; the w-node stresses average Nu_t with the level above, which is
// recomputed here from that level's strain rates
template <typename T>
__global__ void k_nut_tau(const T *__restrict__ S11, const T *__restrict__ S12, const T *__restrict__ S13,
                          const T *__restrict__ S22, const T *__restrict__ S23, const T *__restrict__ S33,
                          const T *__restrict__ Cs, T *__restrict__ Nu_t, T *__restrict__ dissip,
                          T *__restrict__ magS, T *__restrict__ txx, T *__restrict__ txy, T *__restrict__ txz,
                          T *__restrict__ tyy, T *__restrict__ tyz, T *__restrict__ tzz, Geo g, int model,
                          double delta, double Co, int nnn, double vonk, double dz) {
  long i = gtid();
  if (i >= g.np * g.nz) return;
  int k = (int)(i / g.np), x = (int)(i % g.ld);
  T nu = T(0);
  if (x < g.nx) {
    T s, c, l2;
    nu = nut_point<T>(S11, S12, S13, S22, S23, S33, Cs, i, k, model, delta, Co, nnn, vonk, dz, s, c, l2);
    Nu_t[i] = nu;
    dissip[i] = -(s * s * s) * c * l2;
    magS[i] = s;
  }
  if (k == g.nz - 1) {
    txz[i] = T(0);
    tyz[i] = T(0);
    return;
  }
  if (x >= g.nx) return;
  const T m2 = T(-2), h = T(0.5);
  if (k == 0) {
    txx[i] = m2 * nu * S11[i];
    txy[i] = m2 * nu * S12[i];
    tyy[i] = m2 * nu * S22[i];
    tzz[i] = m2 * nu * S33[i];
  } else {
    long p = i + g.np;
    T s1, c1, l21;
    const T nu1 = nut_point<T>(S11, S12, S13, S22, S23, S33, Cs, p, k + 1, model, delta, Co, nnn, vonk, dz, s1,
                               c1, l21);
    T c = h * (nu + nu1);
    txx[i] = m2 * c * h * (S11[i] + S11[p]);
    txy[i] = m2 * c * h * (S12[i] + S12[p]);
    tyy[i] = m2 * c * h * (S22[i] + S22[p]);
    tzz[i] = m2 * c * h * (S33[i] + S33[p]);
    txz[i] = m2 * nu * S13[i];
    tyz[i] = m2 * nu * S23[i];
  }
}

// interpolag_Sdep.f90:62-365: backtrace x - u dt and trilinear interpolation of the four
// Lagrangian averages from the periodic (x, y) / clamped (z) copy of F at point
// (x < nx, y, k); i = its flat index.
// UNIT: cs == 1 (cs_count = 1), where the divisions by cs are the identity and are skipped.
// sx, sy, sz = dt cs / dx, dt cs / dy, dt cs / dz (the backtrace scale factors).
template <typename T, bool UNIT = false>
__device__ __forceinline__ void interp_point(const double *const (&F)[4], double (&out)[4], const T *__restrict__ ul,
                                             const T *__restrict__ vl, const T *__restrict__ wl, Geo g, long i,
                                             int x, int y, int k, double cs, double sx, double sy, double sz) {
  const int r = (int)(i - (long)k * g.np);
  auto dcs = [&](double v) { return UNIT ? v : v / cs; };
  // :204-258 lag velocities on w-nodes (u_lag at jz=1 stays on the uvp node)
  double ut = dcs((double)ul[i]), vt = dcs((double)vl[i]);
  double ulag, vlag, wlag;
  if (k == 0) { ulag = ut; vlag = vt; }
  else {
    ulag = 0.5 * (ut + dcs((double)ul[i - g.np]));
    vlag = 0.5 * (vt + dcs((double)vl[i - g.np]));
  }
  double xp = -ulag * sx;                             // :263-265
  double yp = -vlag * sy;
  double zp;
  if (k <= 1) {                                        // :281-291
    long i2 = (long)g.np + r;
    double w2 = dcs((double)wl[i2]);
    double z2 = -w2 * sz;
    z2 = z2 + fmin(z2, 0.0);
    z2 = fmin(1.0, z2);
    z2 = fmax(-1.0, z2);
    zp = z2;
  } else {
    wlag = dcs((double)wl[i]);
    zp = -wlag * sz;
  }
  int addx = signbit(xp) ? -1 : 1, addy = signbit(yp) ? -1 : 1, addz = signbit(zp) ? -1 : 1;
  double cxp = fabs(xp), cyp = fabs(yp), czp = fabs(zp);
  double fx = 1.0 - cxp, fy = 1.0 - cyp, fz = 1.0 - czp;
  // FF(i', j', k') = F(wrap(i'-1), wrap(j'-1), clamp(k'-1)); point (x,y,k) -> FF(x+1, y+1, k+1)
  int xa = x + addx, ya = y + addy, ka = k + addz;
  xa = (xa + g.nx) % g.nx;
  ya = (ya + g.ny) % g.ny;
  ka = ka < 0 ? 0 : (ka > g.nz - 1 ? g.nz - 1 : ka);
  long o00 = (long)k * g.np + (long)y * g.ld;
  long o0a = (long)ka * g.np + (long)y * g.ld;
  long oa0 = (long)k * g.np + (long)ya * g.ld;
  long oaa = (long)ka * g.np + (long)ya * g.ld;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double *f = F[q];
    out[q] = fx * fy * (f[o00 + x] * fz + f[o0a + x] * czp)
           + fx * cyp * (f[oa0 + x] * fz + f[oaa + x] * czp)
           + cxp * fy * (f[o00 + xa] * fz + f[o0a + xa] * czp)
           + cxp * cyp * (f[oa0 + xa] * fz + f[oaa + xa] * czp);
  }
}

// the whole volume: reads the old F, writes the interpolated values to O; columns
// >= nx keep their old values
template <typename T>
__global__ void k_interpolag(const double *__restrict__ F0, const double *__restrict__ F1,
                             const double *__restrict__ F2, const double *__restrict__ F3,
                             double *__restrict__ I0, double *__restrict__ I1, double *__restrict__ I2,
                             double *__restrict__ I3, const T *__restrict__ ul, const T *__restrict__ vl,
                             const T *__restrict__ wl, Geo g, double cs, double dt, double dx,
                             double dy, double dz) {
  long i = gtid();
  if (i >= g.np * g.nz) return;
  int k = (int)(i / g.np);
  int r = (int)(i % g.np);
  int y = r / g.ld, x = r % g.ld;
  if (x >= g.nx) {
    I0[i] = F0[i]; I1[i] = F1[i]; I2[i] = F2[i]; I3[i] = F3[i];
    return;
  }
  const double *const F[4] = {F0, F1, F2, F3};
  double v[4];
  interp_point<T>(F, v, ul, vl, wl, g, i, x, y, k, cs, dt * cs / dx, dt * cs / dy, dt * cs / dz);
  I0[i] = v[0]; I1[i] = v[1]; I2[i] = v[2]; I3[i] = v[3];
}

// x^(-1/8) for the relaxation time scale (lagrange_Sdep.f90:85 powcoeff = -1/8, a
// constant) as three square roots
__device__ __forceinline__ double pow_m8(double x) { return 1.0 / sqrt(sqrt(sqrt(x))); }

__device__ __forceinline__ double r4max(double a, float b) {  // (double) max(real(a), b)
  return (double)fmaxf((float)a, b);
}

// lagrange_Sdep.f90:157-434 after the filters: Germano contractions, Lagrangian
// relaxation (binary64) with the reference's REAL(4) conversions when casts != 0,
// scale-dependence beta, Cs^2.  bar[q] = 2-delta filtered field q, hat[q] = 4-delta.
// (powcoeff and bexp are the reference's constants -1/8 and 1, checked by the launcher:
// relax_point uses them as such)
struct LagParams {
  double cst, tf1_2, tf2_2, opftdelta, powcoeff, lagran_dt, bexp, tiny, bclip_min;
  int casts, init_lm, init_qn;
};

// Germano contractions at one point from the 21 fields filtered at one width
// (lagrange_Sdep.f90:257-299 for 2 delta, :300-326 for 4 delta): V(q) is filtered field
// q, tf2 the squared filter-width ratio.  Returns L.M (or Q.N) and M.M (or N.N).
template <typename T, typename Vf>
__device__ __forceinline__ void contract_point(Vf V, double cst_d, double tf2, double &lm, double &mm) {
  T ub = V(0), vb = V(1), wb = V(2);
  T L[6];
  L[0] = V(3) - ub * ub; L[1] = V(4) - ub * vb; L[2] = V(5) - ub * wb;
  L[3] = V(6) - vb * vb; L[4] = V(7) - vb * wb; L[5] = V(8) - wb * wb;
  T Sb[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) Sb[q] = V(9 + q);
  T sbar = smag(Sb[0], Sb[1], Sb[2], Sb[3], Sb[4], Sb[5]);
  const T cst = (T)cst_d, t1 = (T)tf2;
  T M[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) M[q] = cst * (V(15 + q) - t1 * sbar * Sb[q]);
  lm = L[0] * M[0] + L[3] * M[3] + L[5] * M[5] + T(2) * (L[1] * M[1] + L[2] * M[2] + L[4] * M[4]);
  mm = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + T(2) * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
}

// lagrange_Sdep.f90:327-434: relaxation of the four averages (binary64, with the
// reference's REAL(4) conversions when casts != 0), Cs^2 and beta.  fi: the interpolated
// (or, in the padding columns, old) averages on entry, the new ones on return.
template <typename T>
__device__ __forceinline__ void relax_point(double LM, double MM, double QN, double NN, double (&fi)[4], bool padc,
                                            bool top, const LagParams &P, T &cs_out, T &beta_out) {
  double flm = fi[0], fmm = fi[1], fqn = fi[2], fnn = fi[3];
  const double tiny = P.tiny;
  if (P.init_lm) { fmm = MM; flm = 0.0; if (padc) { fmm = 1.0; flm = 1.0; } }
  double Tn, epsi;
  if (P.casts) {
    Tn = (double)fmaxf((float)(flm * fmm), 1e-24f);
    Tn = P.opftdelta * pow_m8(Tn);
    Tn = (double)fmaxf(1e-24f, (float)Tn);
  } else {
    Tn = fmax(flm * fmm, tiny);
    Tn = P.opftdelta * pow_m8(Tn);
    Tn = fmax(tiny, Tn);
  }
  epsi = P.lagran_dt / (Tn + P.lagran_dt);  // dum / (1 + dum), dum = lagran_dt / Tn
  flm = (epsi * LM + (1.0 - epsi) * flm);
  fmm = (epsi * MM + (1.0 - epsi) * fmm);
  flm = P.casts ? r4max(flm, 1e-24f) : fmax(tiny, flm);
  double cs2d = flm / fmm;
  if (padc) cs2d = tiny;
  cs2d = P.casts ? (double)(float)fmax(tiny, (double)(float)cs2d) : fmax(tiny, cs2d);
  if (P.init_qn) { fnn = NN; fqn = 0.03 * NN; if (padc) { fnn = 1.0; fqn = 1.0; } }
  if (P.casts) {
    Tn = (double)fmaxf((float)(fqn * fnn), 1e-24f);
    Tn = P.opftdelta * pow_m8(Tn);
    Tn = (double)fmaxf(1e-24f, (float)Tn);
  } else {
    Tn = fmax(fqn * fnn, tiny);
    Tn = P.opftdelta * pow_m8(Tn);
    Tn = fmax(tiny, Tn);
  }
  epsi = P.lagran_dt / (Tn + P.lagran_dt);  // dum / (1 + dum), dum = lagran_dt / Tn
  fqn = (epsi * QN + (1.0 - epsi) * fqn);
  fnn = (epsi * NN + (1.0 - epsi) * fnn);
  fqn = P.casts ? r4max(fqn, 1e-24f) : fmax(tiny, fqn);
  double cs4d = fqn / fnn;
  if (padc) cs4d = tiny;
  cs4d = P.casts ? (double)(float)fmax(tiny, (double)(float)cs4d) : fmax(tiny, cs4d);
  // the exponent log(tf1)/(log(tf2)-log(tf1)) (lagrange_Sdep.f90:410) is exactly 1 for the
  // reference's widths tf1 = 2, tf2 = 4 (:77-78), and pow(x, 1) == x
  double b = cs4d / cs2d;
  if (top) b = 1.0;
  double bclip = P.casts ? (double)fmaxf((float)b, (float)P.bclip_min) : fmax(b, P.bclip_min);
  double csv = cs2d / bclip;
  if (padc) csv = tiny;
  csv = P.casts ? (double)fmaxf(1e-24f, (float)csv) : fmax(tiny, csv);
  fi[0] = flm; fi[1] = fmm; fi[2] = fqn; fi[3] = fnn;
  cs_out = (T)csv;
  beta_out = (T)b;
}

// main.f90:433-439 (+458-460) and 716-728 in one pass: RHS = -conv - divt (+ beta_pcon),
// dudt = -divt (+ mean_p_force on x), then the AB2 predictor (binary64 velocities, the
// update formed in binary64).  first: RHS_f = RHS (main.f90:577-584) for the levels
// written here.
template <typename T>
__global__ void k_rhs_ab2(T *__restrict__ Rx, T *__restrict__ Ry, T *__restrict__ Rz, const T *__restrict__ dx_,
                          const T *__restrict__ dy_, const T *__restrict__ dz_, T *__restrict__ ax,
                          T *__restrict__ ay, T *__restrict__ az, const T *__restrict__ beta_pcon,
                          T *__restrict__ Fx, T *__restrict__ Fy, T *__restrict__ Fz, int first,
                          double *__restrict__ u, double *__restrict__ v, double *__restrict__ w, long n, double dt,
                          double force, T mpf, const T *__restrict__ beta_scal) {
  long i = gtid();
  if (i >= n) return;
  const T rx = -Rx[i] - dx_[i];
  const T ry = -Ry[i] - dy_[i];
  T rz = -Rz[i] - dz_[i];
  if (beta_scal) rz = rz + beta_scal[i];  // main.f90:450-453
  if (beta_pcon) rz = rz + beta_pcon[i];
  Rx[i] = rx;
  Ry[i] = ry;
  Rz[i] = rz;
  if (ax) {  // dudt, dvdt, dwdt: formed when droplets read them (null otherwise)
    ax[i] = -dx_[i] + mpf;
    ay[i] = -dy_[i];
    az[i] = -dz_[i];
  }
  T fx, fy, fz;
  if (first) {
    fx = rx; fy = ry; fz = rz;
    Fx[i] = rx; Fy[i] = ry; Fz[i] = rz;
  } else {
    fx = Fx[i]; fy = Fy[i]; fz = Fz[i];
  }
  const double t1 = 1.5, t2 = 1.0 - 1.5;
  u[i] = u[i] + dt * (t1 * (double)rx + t2 * (double)fx + force);
  v[i] = v[i] + dt * (t1 * (double)ry + t2 * (double)fy);
  w[i] = w[i] + dt * (t1 * (double)rz + t2 * (double)fz);
}

// energy.f90:1-47: one block per level k < nz-1 sums ke = (u^2 + v^2 + w_uv^2)/2 over
// the level in binary64 (per-thread partial sums in a fixed order, then a fixed LDS tree)
static __global__ __launch_bounds__(256) void k_level_ke(const double *__restrict__ u, const double *__restrict__ v,
                                                  const double *__restrict__ w, Geo g, double *__restrict__ out) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  const long base = (long)k * g.np;
  double s = 0.0;
  for (long q = threadIdx.x; q < (long)g.nx * g.ny; q += blockDim.x) {
    const long i = base + (q / g.nx) * g.ld + q % g.nx;
    const double tw = 0.5 * (w[i] + w[i + g.np]);
    s += 0.5 * (u[i] * u[i] + v[i] * v[i] + tw * tw);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// io_crossflow_mod.f90:2647-2799 tavg_compute on levels 1..nz (0-based k < nz): acc[q] +=
// tavg_dt * quantity q (the accumulator struct's order: u, v, w_uv, u_w, v_w, w, u2, v2,
// w2, uv, uw, vw, txx, tyy, tzz, txy, txz, tyz, p - ke, dudx, dudy, dudz_uv, dvdx, dvdy,
// dvdz_uv, dwdx_uv, dwdy_uv, dwdz, and with droplets tot_pc, tot_pc2, then per bin pcon,
// pcon2 as volume fractions).  interp_to_uv_grid (functions.f90:53-128) repeats level
// nz-2 at the top; interp_to_w_grid (:131-187) with u_w, v_w zero at the bottom and top
// levels (:2688-2691).  p has levels 0..nz: its level k+1 is the uv level k.
template <typename T>
__global__ void k_tavg(double *__restrict__ acc, const double *__restrict__ u, const double *__restrict__ v,
                       const double *__restrict__ w, const double *__restrict__ p, const T *__restrict__ txx,
                       const T *__restrict__ tyy, const T *__restrict__ tzz, const T *__restrict__ txy,
                       const T *__restrict__ txz, const T *__restrict__ tyz, const T *__restrict__ dudx,
                       const T *__restrict__ dudy, const T *__restrict__ dudz, const T *__restrict__ dvdx,
                       const T *__restrict__ dvdy, const T *__restrict__ dvdz, const T *__restrict__ dwdx,
                       const T *__restrict__ dwdy, const T *__restrict__ dwdz, const double *__restrict__ PC,
                       const double *__restrict__ pvol, int npc, Geo g, double dt) {
  const long nv = g.np * g.nz;
  const long i = gtid();
  if (i >= nv) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (x >= g.nx) return;
  const long iu = (k < g.nz - 1) ? i : i - g.np;  // lower node of the uv interpolation
  auto uvg = [&](const T *f) { return 0.5 * ((double)f[iu] + (double)f[iu + g.np]); };
  const double up = u[i], vp = v[i], wp = w[i];
  const double w_uv = 0.5 * (w[iu] + w[iu + g.np]);
  const bool edge = k == 0 || k == g.nz - 1;
  const double u_w = edge ? 0.0 : 0.5 * (u[i - g.np] + up);
  const double v_w = edge ? 0.0 : 0.5 * (v[i - g.np] + vp);
  const double q[28] = {up, vp, w_uv, u_w, v_w, wp, up * up, vp * vp, wp * wp, up * vp, u_w * wp, v_w * wp,
                        (double)txx[i], (double)tyy[i], (double)tzz[i], (double)txy[i], (double)txz[i],
                        (double)tyz[i], p[i + g.np] - 0.5 * (up * up + w_uv * w_uv + vp * vp),
                        (double)dudx[i], (double)dudy[i], uvg(dudz), (double)dvdx[i], (double)dvdy[i], uvg(dvdz),
                        uvg(dwdx), uvg(dwdy), (double)dwdz[i]};
#pragma unroll
  for (int m = 0; m < 28; ++m) acc[(long)m * nv + i] += q[m] * dt;
  if (npc > 0) {
    double tot = 0.0;
    for (int b = 0; b < npc; ++b) {
      const double pc = PC[(long)b * nv + i] * pvol[b];
      tot = tot + pc;
      acc[(long)(30 + 2 * b) * nv + i] += pc * dt;
      acc[(long)(31 + 2 * b) * nv + i] += pc * pc * dt;
    }
    acc[28 * nv + i] += tot * dt;
    acc[29 * nv + i] += tot * tot * dt;
  }
}

// main.f90:789-803 (dpdz from p in binary64, RHS -= grad p, dudt -= grad p) and the
// velocity correction of forcing.f90:145-191 in one pass
template <typename T>
__global__ void k_dpdz_project(const double *__restrict__ p, T *__restrict__ dpdz, const T *__restrict__ dpdx,
                               const T *__restrict__ dpdy, T *__restrict__ Rx, T *__restrict__ Ry,
                               T *__restrict__ Rz, T *__restrict__ ax, T *__restrict__ ay, T *__restrict__ az,
                               double *__restrict__ u, double *__restrict__ v, double *__restrict__ w, Geo g,
                               double dz, double dt, int *__restrict__ nonfinite) {
  long i = gtid();
  if (i >= g.np * g.nz) return;
  int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (k == g.nz - 1) { dpdz[i] = (T)LP_BOGUS; return; }
  const double dd = (x < g.nx) ? (p[i + g.np] - p[i]) / dz : 0.0;  // binary64 for the correction
  const T d = (T)dd;
  const T px = dpdx[i], py = dpdy[i];
  dpdz[i] = d;
  Rx[i] -= px;
  Ry[i] -= py;
  Rz[i] -= d;
  if (ax) {
    ax[i] -= px;
    ay[i] -= py;
    az[i] -= d;
  }
  if (x >= g.nx) return;
  const double t1 = 1.5;
  const double un = (u[i] + dt * (-t1 * (double)px + 0.0));
  const double vn = (v[i] + dt * (-t1 * (double)py + 0.0));
  u[i] = un;
  v[i] = vn;
  double wn = 0.0;
  if (k >= 1) {
    wn = (w[i] + dt * (-t1 * dd + 0.0));
    w[i] = wn;
  }
  // the step's velocities are all written here: flag a non-finite one for the host
  if (!isfinite(un) || !isfinite(vn) || !isfinite(wn)) atomicOr(nonfinite, 1);
}

// any non-finite value among a[0..n) (columns x >= nx included) raises *flag
template <typename T>
__global__ void k_nonfinite(const T *__restrict__ a, long n, int *__restrict__ flag) {
  bool bad = false;
  for (long i = gtid(); i < n; i += (long)gridDim.x * blockDim.x) bad |= !isfinite(a[i]);
  if (bad) atomicOr(flag, 1);
}

// forcing.f90:145-191 velocity correction
template <typename T>
__global__ void k_project(double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
                          const T *__restrict__ dpdx, const T *__restrict__ dpdy, const T *__restrict__ dpdz, Geo g,
                          double dt) {
  long i = gtid();
  if (i >= g.np * (g.nz - 1)) return;
  int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (x >= g.nx) return;
  const double t1 = 1.5;
  double R = -t1 * (double)dpdx[i];
  u[i] = (u[i] + dt * (R + 0.0));
  R = -t1 * (double)dpdy[i];
  v[i] = (v[i] + dt * (R + 0.0));
  if (k >= 1) {
    R = -t1 * (double)dpdz[i];
    w[i] = (w[i] + dt * (R + 0.0));
  }
}

// forcing.f90:215-292 boundary values (all columns)
template <typename T>
__global__ void k_project_bc(T *__restrict__ u, T *__restrict__ v, T *__restrict__ w, Geo g, int mean_force,
                             int stress_free) {
  long i = gtid();
  if (i >= g.np) return;
  long top = (long)(g.nz - 1) * g.np + i, bel = (long)(g.nz - 2) * g.np + i;
  if (mean_force) { u[top] = u[bel]; v[top] = v[bel]; }
  w[top] = T(0);
  if (stress_free) { u[i] = u[g.np + i]; v[i] = v[g.np + i]; }
  w[i] = T(0);
}

// press_stag_array.f90:149-329 + tridag_array.f90:30-155, one thread per (kx,ky) mode:
// assemble the wall row, interior rows and lid row on the fly, binary64 Thomas sweep,
// gauge-fixed cumulative sum for the mean mode, zero on the Nyquist row and column.
// H levels 1..nz-1 at planes 0..nz-2 (complex TH, pitch lh), bottomw/topw planes.
template <typename TH>
__global__ __launch_bounds__(64) void k_press_tridiag(const cx<TH> *__restrict__ Hx, const cx<TH> *__restrict__ Hy,
                                const cx<TH> *__restrict__ Hz, const cx<TH> *__restrict__ bw,
                                const cx<TH> *__restrict__ tw_, const double *__restrict__ kxt,
                                const double *__restrict__ kyt, cx<double> *__restrict__ ph,
                                double *__restrict__ gam_w, const double *__restrict__ ibet, int lh, int ny,
                                int nz, long pitch, double dz, int *__restrict__ err) {
  // gam_w: the precomputed gam(j) (binary32 contexts, read only) or the scratch the
  // forward sweep fills (binary64 contexts)
  const double *gam = gam_w;
  const long m = gtid();  // mode (kx, ky) = (m % lh, m / lh)
  const long nh = (long)lh * ny;
  if (m >= nh) return;
  const int jx = (int)(m % lh), jy = (int)(m / lh);
  const long o = (long)jy * pitch + jx;  // offset of the mode in a spectral plane
  const long ps = pitch * ny;            // spectral plane (H, bw, tw, ph)
  auto ld_h = [&](const cx<TH> *H, int level) -> cx<double> {  // H(level), level 1..nz
    if (level >= nz) return mk<double>(0.0, 0.0);
    cx<TH> v = H[(long)(level - 1) * ps + o];
    return mk<double>((double)v.x, (double)v.y);
  };
  if (jy == ny / 2 || jx == lh - 1) {
    for (int j = 0; j <= nz; ++j) ph[(long)j * ps + o] = mk<double>(0.0, 0.0);
    return;
  }
  if (m == 0) {  // press_stag_array.f90:303-312
    cx<double> acc = mk<double>(0.0, 0.0);
    ph[0] = acc;
    cx<TH> b0 = bw[0];
    acc = mk<double>(acc.x - dz * (double)b0.x, acc.y - dz * (double)b0.y);
    ph[ps] = acc;
    for (int jz = 2; jz <= nz; ++jz) {
      cx<double> hz = ld_h(Hz, jz);
      acc = mk<double>(acc.x + hz.x * dz, acc.y + hz.y * dz);
      ph[(long)jz * ps] = acc;
    }
    return;
  }
  const double kx = kxt[m], ky = kyt[m];
  const double idz2 = 1.0 / (dz * dz);
  const double bi = -(kx * kx + ky * ky + 2.0 / (dz * dz));
  // row 0 (wall): b = -1, c = 1, r = -dz * bottomw.  The elimination factors gam(j) and
  // 1/bet(j) depend on the mode and the grid only.  Binary32 contexts take them from the
  // tables of k_press_factors (once per context): the forward sweep is then
  // u(j) = (r(j) - a u(j-1)) * (1/bet(j)), without a division in its chain.  Binary64
  // contexts form them on the way with the reference's divisions (tridag_array.f90:60-80),
  // so that their rounding is the reference's.
  constexpr bool kTab = sizeof(TH) == 4;
  double bet = -1.0, cprev = 1.0;
  cx<TH> b0 = bw[o];
  cx<double> r0 = mk<double>(-dz * (double)b0.x, -dz * (double)b0.y);
  cx<double> uprev;
  if constexpr (kTab) {
    const double ib0 = ibet[m];
    uprev = mk<double>(r0.x * ib0, r0.y * ib0);
  } else {
    uprev = mk<double>(r0.x / bet, r0.y / bet);
  }
  ph[o] = uprev;
  bool bad = false;
  // The right-hand sides and the back-substitution inputs do not depend on the
  // recurrence: they are loaded kCh levels at a time, and the loads of the next chunk are
  // issued before the current chunk is swept (two register buffers), so that the latency
  // of one chunk overlaps the recurrence of the other.  Levels beyond the system are
  // loaded from clamped levels and not used.
  constexpr int kCh = sizeof(TH) == 4 ? 16 : 8;
  struct HChunk {
    cx<TH> x[kCh], y[kCh], z[kCh];
    double ib[kCh];
  };
  auto load_h = [&](int j0, HChunk &c) {
#pragma unroll
    for (int q = 0; q < kCh; ++q) {
      const int j = j0 + q;
      const int lx = j < nz ? j : nz - 1, lz = j + 1 < nz ? j + 1 : nz - 1;
      c.x[q] = Hx[(long)(lx - 1) * ps + o];
      c.y[q] = Hy[(long)(lx - 1) * ps + o];
      c.z[q] = Hz[(long)(lz - 1) * ps + o];
      if constexpr (kTab) c.ib[q] = ibet[(long)(j <= nz ? j : nz) * nh + m];
    }
  };
  cx<double> hz_lo = ld_h(Hz, 1);  // Hz(j) for the next row j
  auto sweep_h = [&](int j0, const HChunk &c) {
#pragma unroll
    for (int q = 0; q < kCh; ++q) {
      const int j = j0 + q;
      if (j > nz) break;
      double a, b, cc;
      cx<double> r;
      if (j < nz) {
        a = idz2; b = bi; cc = idz2;
        const bool okz = j + 1 < nz;
        const cx<double> hx = mk<double>((double)c.x[q].x, (double)c.x[q].y);
        const cx<double> hy = mk<double>((double)c.y[q].x, (double)c.y[q].y);
        const cx<double> hz1 = mk<double>(okz ? (double)c.z[q].x : 0.0, okz ? (double)c.z[q].y : 0.0);
        const cx<double> hz0 = hz_lo;
        // i*(kx Hx + ky Hy) + (Hz(j+1) - Hz(j))/dz
        double sx = kx * hx.x + ky * hy.x, sy = kx * hx.y + ky * hy.y;
        r = mk<double>(-sy + (hz1.x - hz0.x) / dz, sx + (hz1.y - hz0.y) / dz);
        hz_lo = hz1;
      } else {
        cx<TH> t = tw_[o];
        a = -1.0; b = 1.0; cc = 0.0;
        r = mk<double>(-(double)t.x * dz, -(double)t.y * dz);
      }
      cx<double> uj;
      if constexpr (kTab) {
        const double ib = c.ib[q];
        if (isinf(ib)) bad = true;  // bet(j) == 0
        uj = mk<double>((r.x - a * uprev.x) * ib, (r.y - a * uprev.y) * ib);
      } else {
        double gm = cprev / bet;
        gam_w[(long)j * nh + m] = gm;
        bet = b - a * gm;
        if (bet == 0.0) bad = true;
        uj = mk<double>((r.x - a * uprev.x) / bet, (r.y - a * uprev.y) / bet);
      }
      ph[(long)j * ps + o] = uj;
      uprev = uj;
      cprev = cc;
    }
  };
  {
    HChunk ca, cb;
    load_h(1, ca);
    for (int j0 = 1; j0 <= nz; j0 += 2 * kCh) {
      load_h(j0 + kCh, cb);
      sweep_h(j0, ca);
      load_h(j0 + 2 * kCh, ca);
      sweep_h(j0 + kCh, cb);
    }
  }
  // back substitution, from level nz-1 down to 0, with the same two-buffer prefetch
  struct BChunk {
    double g[kCh];
    cx<double> u[kCh];
  };
  auto load_b = [&](int j0, BChunk &c) {
#pragma unroll
    for (int q = 0; q < kCh; ++q) {
      const int j = j0 - q >= 0 ? j0 - q : 0;
      c.g[q] = gam[(long)(j + 1) * nh + m];
      c.u[q] = ph[(long)j * ps + o];
    }
  };
  cx<double> unext = uprev;
  auto sweep_b = [&](int j0, const BChunk &c) {
#pragma unroll
    for (int q = 0; q < kCh; ++q) {
      const int j = j0 - q;
      if (j < 0) break;
      cx<double> uj = mk<double>(c.u[q].x - c.g[q] * unext.x, c.u[q].y - c.g[q] * unext.y);
      ph[(long)j * ps + o] = uj;
      unext = uj;
    }
  };
  {
    BChunk ca, cb;
    load_b(nz - 1, ca);
    for (int j0 = nz - 1; j0 >= 0; j0 -= 2 * kCh) {
      load_b(j0 - kCh, cb);
      sweep_b(j0, ca);
      load_b(j0 - 2 * kCh, ca);
      sweep_b(j0 - kCh, cb);
    }
  }
  if (bad) atomicExch(err, 1);
}

// press_stag_array.f90:149-329 / tridag_array.f90:30-155: the elimination factors of each
// mode's system (a = c = 1/dz^2, b = -(kx^2 + ky^2 + 2/dz^2) inside; the wall row b = -1,
// c = 1; the lid row a = -1, b = 1), which depend on the grid only: gam(j) = c(j-1)/bet(j-1)
// and 1/bet(j), bet(j) = b(j) - a(j) gam(j), in [(nz+1)][lh*ny].  Modes the solve skips
// are left alone.
static __global__ void k_press_factors(const double *__restrict__ kxt, const double *__restrict__ kyt,
                                       double *__restrict__ gam, double *__restrict__ ibet, int lh, int ny, int nz,
                                       double dz) {
  const long m = gtid();
  const long nh = (long)lh * ny;
  if (m >= nh) return;
  const int jx = (int)(m % lh), jy = (int)(m / lh);
  if (jy == ny / 2 || jx == lh - 1 || m == 0) return;
  const double kx = kxt[m], ky = kyt[m];
  const double idz2 = 1.0 / (dz * dz);
  const double bi = -(kx * kx + ky * ky + 2.0 / (dz * dz));
  double bet = -1.0, cprev = 1.0;
  ibet[m] = 1.0 / bet;
  for (int j = 1; j <= nz; ++j) {
    const double a = j < nz ? idz2 : -1.0, b = j < nz ? bi : 1.0, cc = j < nz ? idz2 : 0.0;
    const double gm = cprev / bet;
    gam[(long)j * nh + m] = gm;
    bet = b - a * gm;
    ibet[(long)j * nh + m] = 1.0 / bet;
    cprev = cc;
  }
}

// tridag_array.f90:30-155 with explicit a, b, c, r (all binary64); modes skipped as there
static __global__ void k_tridag_arrays(const double *__restrict__ a, const double *__restrict__ b,
                                const double *__restrict__ c, const cx<double> *__restrict__ r,
                                cx<double> *__restrict__ u, double *__restrict__ gam, int lh, int ny, int n,
                                int *__restrict__ err) {
  long m = gtid();
  long nh = (long)lh * ny;
  if (m >= nh) return;
  int jx = (int)(m % lh), jy = (int)(m / lh);
  double bet = b[m];
  cx<double> r0 = r[m];
  u[m] = mk<double>(r0.x / bet, r0.y / bet);
  if (jx < lh - 1 && bet == 0.0) atomicExch(err, 1);
  if (jy == ny / 2 || jx == lh - 1 || m == 0) return;
  for (int j = 1; j < n; ++j) {
    long s = (long)j * nh + m;
    double gm = c[s - nh] / bet;
    gam[s] = gm;
    bet = b[s] - a[s] * gm;
    if (bet == 0.0) atomicExch(err, 1);
    cx<double> rr = r[s], up = u[s - nh];
    u[s] = mk<double>((rr.x - a[s] * up.x) / bet, (rr.y - a[s] * up.y) / bet);
  }
  for (int j = n - 2; j >= 0; --j) {
    long s = (long)j * nh + m;
    cx<double> uj = u[s], un = u[s + nh];
    double gm = gam[s + nh];
    u[s] = mk<double>(uj.x - gm * un.x, uj.y - gm * un.y);
  }
}

// io.f90:1782-1849 (avg_stats): per-level plane sums of NQ quantities over x < nx, one
// block of 256 threads per level k = blockIdx.x (levels 0..nz-2).  Pass 0: u, v, w,
// Cs_opt2, Nu_t, txz.  Pass 1: u'^2, v'^2, w'^2, u'w' with u', v' on the w-node (mean of
// levels k and k-1, each minus its plane mean) and w' = w - <w>; zero at k = 0 (the
// serial branch of :1811-1815).  out[k*10 + q]: the plane means of the statistics.
template <typename T, int PASS>
__global__ __launch_bounds__(256) void k_plane_stats(const double *__restrict__ u, const double *__restrict__ v,
                                                     const double *__restrict__ w, const T *__restrict__ cs,
                                                     const T *__restrict__ nut, const T *__restrict__ txz, Geo g,
                                                     double *__restrict__ out) {
  constexpr int NQ = PASS == 0 ? 6 : 4;
  __shared__ double red[NQ][256];
  const int k = blockIdx.x;
  const long base = (long)k * g.np;
  double m_u = 0.0, m_v = 0.0, m_w = 0.0, m_um = 0.0, m_vm = 0.0;
  if (PASS == 1) {
    m_u = out[k * 10 + 0]; m_v = out[k * 10 + 1]; m_w = out[k * 10 + 2];
    if (k > 0) { m_um = out[(k - 1) * 10 + 0]; m_vm = out[(k - 1) * 10 + 1]; }
  }
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  const int npts = g.nx * g.ny;
  if (PASS == 0 || k > 0) {
    for (int p = threadIdx.x; p < npts; p += 256) {
      const long i = base + (long)(p / g.nx) * g.ld + p % g.nx;
      if (PASS == 0) {
        acc[0] += (double)u[i]; acc[1] += (double)v[i]; acc[2] += (double)w[i];
        acc[3] += (double)cs[i]; acc[4] += (double)nut[i]; acc[NQ - 1] += (double)txz[i];
      } else {
        const double up = 0.5 * ((double)u[i] - m_u + (double)u[i - g.np] - m_um);
        const double vp = 0.5 * ((double)v[i] - m_v + (double)v[i - g.np] - m_vm);
        const double wp = (double)w[i] - m_w;
        acc[0] += up * up; acc[1] += vp * vp; acc[2] += wp * wp; acc[3] += up * wp;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
#pragma unroll
      for (int q = 0; q < NQ; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nxy = (double)npts;
    if (PASS == 0) {
      for (int q = 0; q < 5; ++q) out[k * 10 + q] = red[q][0] / nxy;
      out[k * 10 + 9] = red[5][0] / nxy;
    } else {
      for (int q = 0; q < 4; ++q) out[k * 10 + 5 + q] = red[q][0] / nxy;
    }
  }
}

// io.f90:1643-1679 (spectrum) from the rows' half spectra X (rows already divided by nx;
// complex pairs of T, row pitch `pitch` complex values, ny rows per level): thread m of
// level blockIdx.y sums w_m |X_m|^2 over the rows in order and divides by ny; w_0 = 1/2 on
// the real part (all of X_0 for a real row), w_m = 1 for 0 < m < nh; no Nyquist term.
template <typename T>
__global__ __launch_bounds__(64) void k_spectrum(const T *__restrict__ spec, long pitch, int ny, int nh,
                                                 double *__restrict__ out) {
  const int m = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y;
  if (m >= nh) return;
  const T *p = spec + 2 * ((long)k * ny * pitch + m);
  double acc = 0.0;
  for (int y = 0; y < ny; ++y, p += 2 * pitch) {
    const double re = (double)p[0], im = (double)p[1];
    acc += m == 0 ? 0.5 * re * re : re * re + im * im;
  }
  out[(long)k * nh + m] = acc / (double)ny;
}

// scalars_module2.f90:1336-1379 and rmsdiv.f90:27-36 partial reductions (levels 1..nz-1)
template <typename T>
__global__ void k_reduce_diag(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
                              const T *__restrict__ nut, const T *__restrict__ du, const T *__restrict__ dv,
                              const T *__restrict__ dw, Geo g, double *__restrict__ part) {
  __shared__ double s_u[256], s_n[256], s_d[256];
  long n = g.np * (g.nz - 1);
  double mu = 0.0, mn = -INFINITY, sd = 0.0;
  for (long i = gtid(); i < n; i += (long)gridDim.x * blockDim.x) {
    int x = (int)(i % g.ld);
    if (x >= g.nx) continue;
    double a = u[i], b = v[i], c = w[i];
    mu = fmax(mu, a * a + b * b + c * c);
    mn = fmax(mn, (double)nut[i]);
    sd += fabs((double)du[i] + (double)dv[i] + (double)dw[i]);
  }
  s_u[threadIdx.x] = mu; s_n[threadIdx.x] = mn; s_d[threadIdx.x] = sd;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s_u[threadIdx.x] = fmax(s_u[threadIdx.x], s_u[threadIdx.x + o]);
      s_n[threadIdx.x] = fmax(s_n[threadIdx.x], s_n[threadIdx.x + o]);
      s_d[threadIdx.x] += s_d[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = s_u[0];
    part[3 * blockIdx.x + 1] = s_n[0];
    part[3 * blockIdx.x + 2] = s_d[0];
  }
}

}  // namespace lp
