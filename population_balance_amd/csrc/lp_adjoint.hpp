// lp_adjoint.hpp -- kernels of the discrete adjoint of the frozen-coefficient step
// (main.f90:252-826 with Nu_t and the wall u* of a base run held fixed), binary64.
// Each kernel is the transpose of a forward kernel, written as a gather (every output
// point sums the contributions that reach it), so no atomics are needed.  Levels k are
// 0-based (Fortran jz = k+1); only the physical columns x < nx carry adjoint values.
#pragma once
#include "lp_droplets.hpp"
#include "lp_kernels.hpp"

namespace lp {

struct AdjGeo {
  int nx, ny, nz, ld;
  long np;
};

__device__ __forceinline__ bool adj_phys(const AdjGeo &g, long i, int &k, int &x) {
  k = (int)(i / g.np);
  x = (int)(i % g.ld);
  return x < g.nx;
}

// forcing.f90:145-292 (velocity correction, boundary values) and main.f90:795-803
// (RHS -= grad p), transposed.  au/av/aw: adjoint of the step's u, v, w; rb*: adjoint of
// the RHS it leaves.  Writes the adjoints of u* (us), of dpdx/dpdy/dpdz (ap*), of the
// RHS before the pressure gradient (ar*), and of filt_da's u, v at the top level (af*:
// the top level of u, v is filt_da's value unless the mean forcing copies it).
__global__ void k_adj_project(const double *__restrict__ au, const double *__restrict__ av,
                              const double *__restrict__ aw, const double *__restrict__ rbx,
                              const double *__restrict__ rby, const double *__restrict__ rbz,
                              double *__restrict__ usx, double *__restrict__ usy, double *__restrict__ usz,
                              double *__restrict__ apx, double *__restrict__ apy, double *__restrict__ apz,
                              double *__restrict__ arx, double *__restrict__ ary, double *__restrict__ arz,
                              double *__restrict__ afu, double *__restrict__ afv, AdjGeo g, double dt,
                              int mean_force, int stress_free) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  int k, x;
  const bool ph = adj_phys(g, i, k, x);
  const long r = i - (long)k * g.np;
  // boundary values: u(top) = u(top-1) with the mean forcing, u(0) = u(1) stress-free
  auto bc = [&](const double *a, long j, int kk) -> double {
    double v = a[j];
    if (kk == g.nz - 1) return 0.0;  // the top level is not a velocity the correction writes
    if (mean_force && kk == g.nz - 2) v += a[(long)(g.nz - 1) * g.np + r];
    if (stress_free && kk == 0) v = 0.0;
    if (stress_free && kk == 1) v += a[r];
    return v;
  };
  double ux = 0, uy = 0, uz = 0, px = 0, py = 0, pz = 0, rx = 0, ry = 0, rz = 0, fu = 0, fv = 0;
  if (ph && k <= g.nz - 2) {
    const double a_u = bc(au, i, k), a_v = bc(av, i, k);
    const double a_w = (k >= 1) ? aw[i] : 0.0;  // w(0) = 0 is a boundary value
    ux = a_u; uy = a_v; uz = a_w;
    px = -1.5 * dt * a_u - rbx[i];
    py = -1.5 * dt * a_v - rby[i];
    pz = -1.5 * dt * a_w - rbz[i];
    rx = rbx[i]; ry = rby[i]; rz = rbz[i];
  }
  if (ph && k == g.nz - 1 && !mean_force) { fu = au[i]; fv = av[i]; }
  usx[i] = ux; usy[i] = uy; usz[i] = uz;
  apx[i] = px; apy[i] = py; apz[i] = pz;
  arx[i] = rx; ary[i] = ry; arz[i] = rz;
  afu[i] = fu; afv[i] = fv;
}

// dpdz = (p(k+1) - p(k))/dz on levels 0..nz-2 (main.f90:789-791), transposed: the
// adjoint of p on planes 0..nz-1 (plane nz of p is not an output)
__global__ void k_adj_dpdz(const double *__restrict__ apz, double *__restrict__ ap, AdjGeo g, double idz) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  int k, x;
  double v = 0.0;
  if (adj_phys(g, i, k, x)) {
    if (k >= 1) v += apz[i - g.np] * idz;
    if (k <= g.nz - 2) v -= apz[i] * idz;
  }
  ap[i] = v;
}

// press_stag_array.f90:149-329 with tridag_array.f90, transposed, one thread per mode.
// Input spectra (forward y transforms of the r2c, pitch `pitch` complex per row):
// Sp[j] of the adjoint of p (planes 0..nz-1), Sx[j-1], Sy[j-1] of the adjoints of dpdx,
// dpdy (j = 1..nz-1).  The adjoint of p_hat is W (Sp - i kx Sx - i ky Sy), W = 2 on the
// interior kx (c2r^T).  The transposed system A^T lambda = that is solved (A: wall row
// (-1, 1), interior rows (1/dz^2, bi, 1/dz^2), lid row (-1, 1)); the mean mode is a
// reverse cumulative sum.  Outputs, already multiplied by the halving of the interior kx
// (r2c^T) and the forward scales (cst/(1.5 dt) for H, cst for the wall and lid rows):
// Hx, Hy, Hz planes 0..nz-2 (H levels 1..nz-1), bw, tw one plane each.
__global__ __launch_bounds__(64) void k_press_adj(const cx<double> *__restrict__ Sp, const cx<double> *__restrict__ Sx,
                                                  const cx<double> *__restrict__ Sy, const double *__restrict__ kxt,
                                                  const double *__restrict__ kyt, cx<double> *__restrict__ Hx,
                                                  cx<double> *__restrict__ Hy, cx<double> *__restrict__ Hz,
                                                  cx<double> *__restrict__ bw, cx<double> *__restrict__ tw,
                                                  double *__restrict__ gam, cx<double> *__restrict__ lam, int lh,
                                                  int ny, int nz, long pitch, double dz, double sH, double sB) {
  const long m = gtid();
  const long nh = (long)lh * ny;
  if (m >= nh) return;
  const int jx = (int)(m % lh), jy = (int)(m / lh);
  const long o = (long)jy * pitch + jx;
  const long ps = pitch * ny;
  const cx<double> z0 = mk<double>(0.0, 0.0);
  if (jy == ny / 2 || jx == lh - 1) {  // Nyquist modes: p_hat is zero there
    for (int j = 0; j < nz - 1; ++j) { Hx[j * ps + o] = z0; Hy[j * ps + o] = z0; Hz[j * ps + o] = z0; }
    bw[o] = z0; tw[o] = z0;
    return;
  }
  const double W = (jx == 0) ? 1.0 : 2.0, h = (jx == 0) ? 1.0 : 0.5;
  const double kx = kxt[m], ky = kyt[m];
  auto rhs = [&](int j) -> cx<double> {  // adjoint of p_hat(j), j = 0..nz
    if (j >= nz) return z0;
    cx<double> v = Sp[(long)j * ps + o];
    if (j >= 1) {
      const cx<double> a = Sx[(long)(j - 1) * ps + o], b = Sy[(long)(j - 1) * ps + o];
      // -i kx a - i ky b
      v = mk<double>(v.x + kx * a.y + ky * b.y, v.y - kx * a.x - ky * b.x);
    }
    return mk<double>(W * v.x, W * v.y);
  };
  if (m == 0) {  // mean mode :303-312: p(1) = -dz bw, p(j) = p(j-1) + dz Hz(j)
    cx<double> acc = z0;
    for (int j = nz; j >= 2; --j) {
      const cx<double> r = rhs(j);
      acc = mk<double>(acc.x + r.x, acc.y + r.y);
      if (j <= nz - 1) Hz[(long)(j - 1) * ps + o] = mk<double>(h * sH * dz * acc.x, h * sH * dz * acc.y);
    }
    const cx<double> r1 = rhs(1);
    acc = mk<double>(acc.x + r1.x, acc.y + r1.y);
    bw[o] = mk<double>(h * sB * -dz * acc.x, h * sB * -dz * acc.y);
    tw[o] = z0;
    Hx[o] = z0; Hy[o] = z0;
    for (int j = 1; j < nz - 1; ++j) { Hx[j * ps + o] = z0; Hy[j * ps + o] = z0; }
    Hz[o] = z0;  // H level 1 of the mean mode is not read by the forward
    return;
  }
  const double idz2 = 1.0 / (dz * dz);
  const double bi = -(kx * kx + ky * ky + 2.0 / (dz * dz));
  // A^T: a'_j = c_{j-1}, b'_j = b_j, c'_j = a_{j+1}
  auto ap_ = [&](int j) { return j == 0 ? 0.0 : (j - 1 == 0 ? 1.0 : idz2); };
  auto bp_ = [&](int j) { return j == 0 ? -1.0 : (j == nz ? 1.0 : bi); };
  auto cp_ = [&](int j) { return j == nz ? 0.0 : (j + 1 == nz ? -1.0 : idz2); };
  double bet = bp_(0);
  cx<double> r0 = rhs(0);
  cx<double> up = mk<double>(r0.x / bet, r0.y / bet);
  lam[m] = up;
  for (int j = 1; j <= nz; ++j) {
    const double gm = cp_(j - 1) / bet;
    gam[(long)j * nh + m] = gm;
    bet = bp_(j) - ap_(j) * gm;
    const cx<double> r = rhs(j);
    const double a = ap_(j);
    up = mk<double>((r.x - a * up.x) / bet, (r.y - a * up.y) / bet);
    lam[(long)j * nh + m] = up;
  }
  // back substitution from lambda(nz) down; iteration j knows lambda(j) and lambda(j+1)
  cx<double> nxt = up;  // lambda(nz)
  tw[o] = mk<double>(h * sB * -dz * nxt.x, h * sB * -dz * nxt.y);
  for (int j = nz - 1; j >= 0; --j) {
    const cx<double> l0 = lam[(long)j * nh + m];
    const double gm = gam[(long)(j + 1) * nh + m];
    const cx<double> lj = mk<double>(l0.x - gm * nxt.x, l0.y - gm * nxt.y);
    if (j >= 1) {  // row j: i (kx Hx(j) + ky Hy(j)) + (Hz(j+1) - Hz(j))/dz
      const long q = (long)(j - 1) * ps + o;
      Hx[q] = mk<double>(h * sH * kx * lj.y, h * sH * -kx * lj.x);  // -i kx lambda
      Hy[q] = mk<double>(h * sH * ky * lj.y, h * sH * -ky * lj.x);
    }
    if (j <= nz - 2) {  // Hz(j+1): +lambda(j)/dz from row j (j >= 1), -lambda(j+1)/dz from row j+1
      const double ax = (j >= 1 ? lj.x : 0.0) - nxt.x, ay = (j >= 1 ? lj.y : 0.0) - nxt.y;
      Hz[(long)j * ps + o] = mk<double>(h * sH * ax / dz, h * sH * ay / dz);
    }
    if (j == 0) bw[o] = mk<double>(h * sB * -dz * lj.x, h * sB * -dz * lj.y);
    nxt = lj;
  }
}

}  // namespace lp

namespace lp {

// forward intermediates of the adjoint's base state: the wall-level dudz, dvdz of the
// wall law (u / (zl denom), wallstress.f90:176-183 with the frozen u*) and the vorticity
// of convec.f90:100-133 (jz = 1 on the uvp node)
__global__ void k_adj_base(const double *__restrict__ uf, const double *__restrict__ vf,
                           const double *__restrict__ uz, const double *__restrict__ vz, const double *__restrict__ uy, const double *__restrict__ vx,
                           const double *__restrict__ wx, const double *__restrict__ wy, double *__restrict__ om1,
                           double *__restrict__ om2, double *__restrict__ om3, AdjGeo g, double wfac, int wall) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  int k, x;
  if (!adj_phys(g, i, k, x)) { om1[i] = 0; om2[i] = 0; om3[i] = 0; return; }
  const double a_uz = k == 0 ? (wall ? uf[i] * wfac : 0.0) : uz[i];
  const double a_vz = k == 0 ? (wall ? vf[i] * wfac : 0.0) : vz[i];
  const long up = (k == 0) ? g.np : 0;
  const double wyv = k == 0 ? 0.5 * (wy[i] + wy[i + up]) : wy[i];
  const double wxv = k == 0 ? 0.5 * (wx[i] + wx[i + up]) : wx[i];
  om1[i] = wyv - a_vz;
  om2[i] = a_uz - wxv;
  om3[i] = vx[i] - uy[i];
}

// the AB2 predictor (main.f90:716-728), the RHS assembly (:433-439) and the Euler start
// (:577-584), transposed, together with the pressure's contributions: us* (projection)
// + h* (H back-transformed) is the adjoint of u*; wall / lid rows (bwb, twb planes) go to
// divtz.  Outputs the adjoints of filt_da's u, v, w (af*, initialised here), of the
// convective terms (ac*), of the stress divergences (ad*) and of RHS_f (rb*, in place).
__global__ void k_adj_ab2(const double *__restrict__ usx, const double *__restrict__ usy,
                          const double *__restrict__ usz, const double *__restrict__ hx, const double *__restrict__ hy,
                          const double *__restrict__ hz, const double *__restrict__ arx,
                          const double *__restrict__ ary, const double *__restrict__ arz,
                          const double *__restrict__ afu_top, const double *__restrict__ afv_top,
                          const double *__restrict__ bwb, const double *__restrict__ twb, double *__restrict__ afu,
                          double *__restrict__ afv, double *__restrict__ afw, double *__restrict__ acx,
                          double *__restrict__ acy, double *__restrict__ acz, double *__restrict__ adx,
                          double *__restrict__ ady, double *__restrict__ adz, double *__restrict__ rbx,
                          double *__restrict__ rby, double *__restrict__ rbz, AdjGeo g, double dt, int first) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  int k, x;
  const bool ph = adj_phys(g, i, k, x);
  const long r = i - (long)k * g.np;
  double fu = 0, fv = 0, fw = 0, cx_ = 0, cy_ = 0, cz_ = 0, dx_ = 0, dy_ = 0, dz_ = 0, ox = 0, oy = 0, oz = 0;
  if (ph && k <= g.nz - 2) {
    const double sx = usx[i] + hx[i], sy = usy[i] + hy[i], sz = usz[i] + hz[i];
    const double tr = first ? 1.0 : 1.5;
    fu = sx; fv = sy; fw = sz;
    const double rx = arx[i] + tr * dt * sx, ry = ary[i] + tr * dt * sy, rz = arz[i] + tr * dt * sz;
    ox = first ? 0.0 : -0.5 * dt * sx;
    oy = first ? 0.0 : -0.5 * dt * sy;
    oz = first ? 0.0 : -0.5 * dt * sz;
    cx_ = -rx; dx_ = -rx;
    cy_ = -ry; dy_ = -ry;
    cz_ = k >= 1 ? -rz : 0.0;
    dz_ = -rz;
  }
  if (ph && k == 0) dz_ += bwb[r];
  if (ph && k == g.nz - 1) {
    dz_ += twb[r];
    fu = afu_top[i];
    fv = afv_top[i];
  }
  afu[i] = fu; afv[i] = fv; afw[i] = fw;
  acx[i] = cx_; acy[i] = cy_; acz[i] = cz_;
  adx[i] = dx_; ady[i] = dy_; adz[i] = dz_;
  rbx[i] = ox; rby[i] = oy; rbz[i] = oz;
}

// convec.f90:138-279 products on the 3/2 grid, transposed (gather over the levels whose
// products read level k).  p*: adjoint of the products (times the truncation's transpose
// scale tt); U, O: base velocities and vorticity on the 3/2 grid (nz levels, npb points).
__global__ void k_adj_products(const double *__restrict__ px, const double *__restrict__ py,
                               const double *__restrict__ pz, const double *__restrict__ U1,
                               const double *__restrict__ U2, const double *__restrict__ U3,
                               const double *__restrict__ O1, const double *__restrict__ O2,
                               const double *__restrict__ O3, double *__restrict__ aU1, double *__restrict__ aU2,
                               double *__restrict__ aU3, double *__restrict__ aO1, double *__restrict__ aO2,
                               double *__restrict__ aO3, long npb, int nz, double tt) {
  const long a = gtid();
  if (a >= npb * nz) return;
  const int k = (int)(a / npb);
  const long m = a - npb, b = a + npb;
  auto P = [&](const double *p, long j, int kk, int lo) -> double {  // tt p(kk), kk in lo..nz-2
    return (kk >= lo && kk <= nz - 2) ? tt * p[j] : 0.0;
  };
  const double pxk = P(px, a, k, 0), pyk = P(py, a, k, 0), pzk = P(pz, a, k, 1);
  const double pxm = P(px, m, k - 1, 0), pym = P(py, m, k - 1, 0);
  const double pzp = P(pz, b, k + 1, 1);
  const double u1 = U1[a], u2 = U2[a], u3 = U3[a], o1 = O1[a], o2 = O2[a], o3 = O3[a];
  const double u1m = k >= 1 ? U1[m] : 0.0, u2m = k >= 1 ? U2[m] : 0.0;
  const double o1p = k + 1 <= nz - 1 ? O1[b] : 0.0, o2p = k + 1 <= nz - 1 ? O2[b] : 0.0;
  const bool in = k >= 1;  // the [k >= 1] terms of px, py at level k
  aU1[a] = o3 * pyk - 0.5 * o2 * pzk - 0.5 * o2p * pzp;
  aU2[a] = -o3 * pxk + 0.5 * o1 * pzk + 0.5 * o1p * pzp;
  aU3[a] = 0.5 * o2 * pxm + (in ? 0.5 * o2 * pxk : 0.0) - 0.5 * o1 * pym - (in ? 0.5 * o1 * pyk : 0.0);
  aO1[a] = -0.5 * u3 * pym - (in ? 0.5 * u3 * pyk : 0.0) + 0.5 * (u2 + u2m) * pzk;
  aO2[a] = 0.5 * u3 * pxm + (in ? 0.5 * u3 * pxk : 0.0) - 0.5 * (u1 + u1m) * pzk;
  aO3[a] = -u2 * pxk + u1 * pyk;
}

// divstress_uv / divstress_w (z parts) and the stresses with the frozen Nu_t
// (sgs_stag.f90:254-325), transposed: the adjoints of S11..S33 from the adjoints of the
// divergences (ad*) and of their x / y derivative parts (d*: Dx ad, Dy ad, already
// computed; the derivatives are antisymmetric so they enter with a minus sign).
// qw: the wall-law coefficient times the adjoint of txz(1), tyz(1) is written to
// wtx / wty (plane) for the test filter.
__global__ void k_adj_tau(const double *__restrict__ adx, const double *__restrict__ ady,
                          const double *__restrict__ adz, const double *__restrict__ dxa,
                          const double *__restrict__ dya, const double *__restrict__ dxb,
                          const double *__restrict__ dyb, const double *__restrict__ dxc,
                          const double *__restrict__ dyc, const double *__restrict__ nut,
                          const double *__restrict__ ust, double *__restrict__ aS11, double *__restrict__ aS12,
                          double *__restrict__ aS13, double *__restrict__ aS22, double *__restrict__ aS23,
                          double *__restrict__ aS33, double *__restrict__ wtx, double *__restrict__ wty, AdjGeo g,
                          double idz, double qfac, int wall) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  int k, x;
  if (!adj_phys(g, i, k, x)) {
    aS11[i] = 0; aS12[i] = 0; aS13[i] = 0; aS22[i] = 0; aS23[i] = 0; aS33[i] = 0;
    return;
  }
  const int nz = g.nz;
  const long np = g.np;
  // adjoints of the stresses at level kk (this column)
  auto atxx = [&](int kk) { return -dxa[i + (long)(kk - k) * np]; };
  auto atxy = [&](int kk) { const long j = i + (long)(kk - k) * np; return -dya[j] - dxb[j]; };
  auto atyy = [&](int kk) { return -dyb[i + (long)(kk - k) * np]; };
  auto atxz = [&](int kk) {
    const long j = i + (long)(kk - k) * np;
    double v = -dxc[j];
    if (kk - 1 >= 0 && kk - 1 <= nz - 2) v += adx[j - np] * idz;
    if (kk <= nz - 2) v -= adx[j] * idz;
    return v;
  };
  auto atyz = [&](int kk) {
    const long j = i + (long)(kk - k) * np;
    double v = -dyc[j];
    if (kk - 1 >= 0 && kk - 1 <= nz - 2) v += ady[j - np] * idz;
    if (kk <= nz - 2) v -= ady[j] * idz;
    return v;
  };
  auto atzz = [&](int kk) {
    const long j = i + (long)(kk - k) * np;
    double v = 0.0;
    if (kk >= 1 && kk <= nz - 2) v += adz[j] * idz;
    if (kk + 1 >= 1 && kk + 1 <= nz - 2) v -= adz[j + np] * idz;
    return v;
  };
  auto cc = [&](int kk) { const long j = i + (long)(kk - k) * np; return 0.5 * (nut[j] + nut[j + np]); };
  double s11 = 0, s12 = 0, s13 = 0, s22 = 0, s23 = 0, s33 = 0;
  if (k == 0) {
    const double n0 = nut[i];
    s11 = -2.0 * n0 * atxx(0); s12 = -2.0 * n0 * atxy(0); s22 = -2.0 * n0 * atyy(0); s33 = -2.0 * n0 * atzz(0);
    if (wall) {  // txz(1) = q G(u(1)), q = -u* kappa / denom (frozen)
      wtx[i] = -ust[i] * qfac * atxz(0);
      wty[i] = -ust[i] * qfac * atyz(0);
    }
  } else {
    if (k <= nz - 2) {
      const double c0 = cc(k);
      s11 = -c0 * atxx(k); s12 = -c0 * atxy(k); s22 = -c0 * atyy(k); s33 = -c0 * atzz(k);
      s13 = -2.0 * nut[i] * atxz(k);
      s23 = -2.0 * nut[i] * atyz(k);
    }
    if (k - 1 >= 1 && k - 1 <= nz - 2) {
      const double c1 = cc(k - 1);
      s11 += -c1 * atxx(k - 1); s12 += -c1 * atxy(k - 1); s22 += -c1 * atyy(k - 1); s33 += -c1 * atzz(k - 1);
    }
  }
  aS11[i] = s11; aS12[i] = s12; aS13[i] = s13; aS22[i] = s22; aS23[i] = s23; aS33[i] = s33;
}

// the strain rates (sgs_stag.f90:49-125), the vorticity (convec.f90:100-133) and the z
// derivatives (ddz_uv, ddz_w, the wall law's dudz) transposed: the adjoints of filt_da's
// u, v, w gain the z-derivative parts here (af*, in place), and those of the x and y
// derivatives are written (a*x, a*y).  ao*: adjoint of the vorticity.
__global__ void k_adj_strain(const double *__restrict__ aS11, const double *__restrict__ aS12,
                             const double *__restrict__ aS13, const double *__restrict__ aS22,
                             const double *__restrict__ aS23, const double *__restrict__ aS33,
                             const double *__restrict__ ao1, const double *__restrict__ ao2,
                             const double *__restrict__ ao3, double *__restrict__ aux, double *__restrict__ auy,
                             double *__restrict__ avx, double *__restrict__ avy, double *__restrict__ awx,
                             double *__restrict__ awy, double *__restrict__ afu, double *__restrict__ afv,
                             double *__restrict__ afw, AdjGeo g, double idz, double wfac, int wall) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  int k, x;
  if (!adj_phys(g, i, k, x)) {
    aux[i] = 0; auy[i] = 0; avx[i] = 0; avy[i] = 0; awx[i] = 0; awy[i] = 0;
    return;
  }
  const int nz = g.nz;
  const long np = g.np, r = i - (long)k * np;
  auto at = [&](const double *a, int kk) { return a[r + (long)kk * np]; };
  const bool hasp = k + 1 <= nz - 1;
  // x and y derivatives
  aux[i] = (k == 0 ? at(aS11, 0) : 0.5 * at(aS11, k)) + (hasp ? 0.5 * at(aS11, k + 1) : 0.0);
  const double s12 = (k == 0 ? 0.5 * at(aS12, 0) : 0.25 * at(aS12, k)) + (hasp ? 0.25 * at(aS12, k + 1) : 0.0);
  auy[i] = s12 - at(ao3, k);
  avx[i] = s12 + at(ao3, k);
  avy[i] = (k == 0 ? at(aS22, 0) : 0.5 * at(aS22, k)) + (hasp ? 0.5 * at(aS22, k + 1) : 0.0);
  awx[i] = (k <= 1 ? 0.25 * at(aS13, 0) - 0.5 * at(ao2, 0) : 0.0) + (k >= 1 ? 0.5 * at(aS13, k) - at(ao2, k) : 0.0);
  awy[i] = (k <= 1 ? 0.25 * at(aS23, 0) + 0.5 * at(ao1, 0) : 0.0) + (k >= 1 ? 0.5 * at(aS23, k) + at(ao1, k) : 0.0);
  // z derivatives: auz, avz at level kk; awz at level kk
  auto auz = [&](int kk) { return (kk == 0 ? 0.5 * at(aS13, 0) : 0.5 * at(aS13, kk)) + at(ao2, kk); };
  auto avz = [&](int kk) { return (kk == 0 ? 0.5 * at(aS23, 0) : 0.5 * at(aS23, kk)) - at(ao1, kk); };
  auto awz = [&](int kk) {
    return (kk == 0 ? at(aS33, 0) : 0.5 * at(aS33, kk)) + (kk + 1 <= nz - 1 ? 0.5 * at(aS33, kk + 1) : 0.0);
  };
  double fu = afu[i], fv = afv[i], fw = afw[i];
  // dudz(k) = (u(k) - u(k-1))/dz on 1..nz-2; dudz(0) = u(0) wfac (wall law)
  if (k >= 1 && k <= nz - 2) { fu += auz(k) * idz; fv += avz(k) * idz; }
  if (k + 1 >= 1 && k + 1 <= nz - 2) { fu -= auz(k + 1) * idz; fv -= avz(k + 1) * idz; }
  if (k == 0 && wall) { fu += auz(0) * wfac; fv += avz(0) * wfac; }
  // dwdz(k) = (w(k+1) - w(k))/dz on 0..nz-2
  if (k >= 1) fw += awz(k - 1) * idz;
  if (k <= nz - 2) fw -= awz(k) * idz;
  afu[i] = fu; afv[i] = fv; afw[i] = fw;
}

// out = a - b - c on the physical columns, zero on the padding (filt_da transposed:
// F(af) - Dx(ax) - Dy(ay)); with wall plane add: af(0) += gw before (the wall-law term)
__global__ void k_adj_combine(double *__restrict__ out, const double *__restrict__ a, const double *__restrict__ b,
                              const double *__restrict__ c, AdjGeo g) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  int k, x;
  out[i] = adj_phys(g, i, k, x) ? a[i] - b[i] - c[i] : 0.0;
}

__global__ void k_add_plane(double *__restrict__ a, const double *__restrict__ b, long n) {
  const long i = gtid();
  if (i < n) a[i] += b[i];
}

}  // namespace lp

namespace lp {

// wallstress.f90:101-185 with the wall u* of a base run (frozen coefficient):
// txz = -(u*^2/|U|) u1 = -(u* kappa/denom) u1, dudz = u/(zl denom); u1, v1 test-filtered
template <typename T>
__global__ void k_wallstress_frozen(const T *__restrict__ u1, const T *__restrict__ v1, const double *__restrict__ u,
                                    const double *__restrict__ v, T *__restrict__ txz, T *__restrict__ tyz,
                                    T *__restrict__ dudz, T *__restrict__ dvdz, T *__restrict__ ust,
                                    const T *__restrict__ frz_ust, Geo g, double denom, double zl, double vonk) {
  long i = gtid();
  if (i >= g.np) return;
  int x = (int)(i % g.ld);
  if (x >= g.nx) return;
  const double q = -(double)frz_ust[i] * vonk / denom;
  txz[i] = (T)(q * (double)u1[i]);
  tyz[i] = (T)(q * (double)v1[i]);
  dudz[i] = (T)(u[i] / (zl * denom));
  dvdz[i] = (T)(v[i] / (zl * denom));
  ust[i] = frz_ust[i];
}

// sgs_stag.f90:254-325, 391-394 with the Nu_t of a base run (frozen coefficient): the
// stresses of k_nut_tau from the strain rates and a given Nu_t (copied to Nu_t)
template <typename T>
__global__ void k_tau_frozen(const T *__restrict__ S11, const T *__restrict__ S12, const T *__restrict__ S13,
                             const T *__restrict__ S22, const T *__restrict__ S23, const T *__restrict__ S33,
                             const T *__restrict__ frz, T *__restrict__ Nu_t, T *__restrict__ txx,
                             T *__restrict__ txy, T *__restrict__ txz, T *__restrict__ tyy, T *__restrict__ tyz,
                             T *__restrict__ tzz, Geo g) {
  long i = gtid();
  if (i >= g.np * g.nz) return;
  int k = (int)(i / g.np), x = (int)(i % g.ld);
  const T nu = frz[i];
  if (x < g.nx) Nu_t[i] = nu;
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
    const long p = i + g.np;
    const T c = h * (nu + frz[p]);
    txx[i] = m2 * c * h * (S11[i] + S11[p]);
    txy[i] = m2 * c * h * (S12[i] + S12[p]);
    tyy[i] = m2 * c * h * (S22[i] + S22[p]);
    tzz[i] = m2 * c * h * (S33[i] + S33[p]);
    txz[i] = m2 * nu * S13[i];
    tyz[i] = m2 * nu * S23[i];
  }
}

// ------------------------------------------------------------------------------------
// droplets (pollen_all_in_one2, scalars_module_crossflow_mod.f90:1494-1786, with
// pollen_RHS_calc2 :1894-4472, calcbeta_pcon :1795-1843, step_pollen2 :4481-4508), all
// bins in one launch per pass, with per-bin face adjoints summed at the end.  Frozen: Kc_t (from Nu_t).  The transport is linear in C on each branch
// of the SMART limiter: its transpose uses the branch the base C selects.  The three
// directions are gathers (a thread per cell adds what the line scan of the fluxes
// transposed would scatter into it, in the scan's order), so no pass needs atomics.
// ------------------------------------------------------------------------------------

// d smart_face / d (cup, cupup, cdn) on the branch smart_face selects
__device__ __forceinline__ void smart_coef(double cup, double cupup, double cdn, double &au, double &auu,
                                           double &ad) {
  au = 1.0; auu = 0.0; ad = 0.0;
  const double den = cup - cupup;
  if (den == 0.0) return;
  const double num = cdn - cup;
  const double a = 2.0 * num, b = 0.75 * num + 0.25 * den, c = 4.0 * den;
  const double pd = den > 0.0 ? fmax(0.0, fmin(fmin(a, b), c)) : fmin(0.0, fmax(fmax(a, b), c));
  if (pd == 0.0) return;                                              // the upwind cell
  if (pd == a) { au = 0.0; ad = 1.0; return; }                        // the downwind cell
  if (pd == b) { au = 0.75; auu = -0.125; ad = 0.375; return; }       // QUICK
  au = 3.0; auu = -2.0;                                               // psi = 4
}

// step_pollen2 and calcbeta_pcon transposed (every point of every bin):
// lb, lr the adjoints of the step's PCon and RHS_PCon (lr becomes that of the RHS_PCon the
// step reads); mu = adjoint of this step's RHS_PCon; ac = adjoint of the old PCon through
// the update and the buoyancy (the transport adds to it)
__global__ void k_adj_drop_update(const double *__restrict__ lb_all, double *__restrict__ lr_all,
                                  const double *__restrict__ arz, const double *__restrict__ vfac, int active,
                                  double *__restrict__ mu_all, double *__restrict__ ac_all, DropGeo g, double dt,
                                  int euler) {
  const long t = gtid();
  if (t >= g.nv * g.npc) return;
  const int b = (int)(t / g.nv);
  const long s = t % g.nv, ob = (long)b * g.nv;
  const double *lb = lb_all + ob;
  double *lr = lr_all + ob, *mu = mu_all + ob, *ac = ac_all + ob;
  const int k = (int)(s / g.np), x = (int)(s % g.ld);
  if (k == g.nz - 1 || x >= g.nx) {
    mu[s] = 0.0; ac[s] = 0.0; lr[s] = 0.0;
    return;
  }
  const double l = lb[s] + (k == g.nz - 2 ? lb[s + g.np] : 0.0);  // the top level copies nz-2
  mu[s] = lr[s] + (euler ? dt : 1.5 * dt) * l;
  ac[s] = l + (active ? vfac[b] * arz[s] : 0.0);
  lr[s] = euler ? 0.0 : -0.5 * dt * l;
}

struct DropAdjP {
  double idx, idy, idz, idx2, idy2, idz2;
  int bc_x, bc_y, settling;
};

// x (Y false) and y (Y true) transport transposed in gather form for one cell of one bin.  A cell adds the terms the line scan of
// the fluxes transposed scatters into it, in the scan's order: the interior faces c-1, c,
// c+1, c+2, then the end faces, the Kc diffusion of cells c-1, c, c+1 (ghosts at the line
// ends), and last the folds of the ghosts' adjoints at the two end cells.  The terms of a
// face are recomputed by each cell that reads them; cell c = f >= 1 writes face f's
// adjoint into auv (the bin's (ld, ny+1, nz) face array), cell 0 or n-1 the end faces.
template <bool Y>
__device__ __forceinline__ double drop_line_cell(double A, int i, int j, int k, int b, const double *__restrict__ C,
                                                 const double *__restrict__ mu, const double *__restrict__ uv0,
                                                 double *__restrict__ auv, const double *__restrict__ kc,
                                                 const DropGeo &g, const DropAdjP &P) {
  const int n = Y ? g.ny : g.nx, c = Y ? j : i, bc = Y ? P.bc_y : P.bc_x;
  const long st = Y ? g.ld : 1, ol = Y ? (long)i : (long)j * g.ld, ob = (long)b * g.nv;
  const double *Cl = C + ob + (long)k * g.np + ol, *Ml = mu + ob + (long)k * g.np + ol, *Kl = kc + (long)k * g.np + ol;
  const double *U = uv0 + (long)k * g.fp + ol;
  double *AU = auv + (long)b * g.fp * g.nz + (long)k * g.fp + ol;
  const double id = Y ? P.idy : P.idx, id2 = Y ? P.idy2 : P.idx2;
  auto Cv = [&](int q) { return Cl[q * st]; };
  auto Mv = [&](int q) { return Ml[q * st]; };
  auto Kv = [&](int q) { return Kl[q * st]; };
  // the cell's stencil, loaded together: C at c-3 .. c+3, mu at c-2 .. c+2, Kc at c-2 .. c+2
  // (periodic), faces c-1 .. c+2; indices off the line are clamped (their values go unused)
  double Cw[7], Mw[5], Kw[5], Uw[4];
#pragma unroll
  for (int d = 0; d < 7; ++d) Cw[d] = Cv(min(max(c + d - 3, 0), n - 1));
#pragma unroll
  for (int d = 0; d < 5; ++d) Mw[d] = Mv(min(max(c + d - 2, 0), n - 1));
#pragma unroll
  for (int d = 0; d < 5; ++d) {
    const int q = c + d - 2;
    Kw[d] = Kv(q < 0 ? q + n : (q >= n ? q - n : q));
  }
#pragma unroll
  for (int d = 0; d < 4; ++d) Uw[d] = U[min(max(c + d - 1, 0), n) * st];
  const double u0 = U[0], Cf = Cv(0), Cn = Cv(n - 1);
  double g0 = 0.0, gL = 0.0, s0 = 0.0, s1 = 0.0;  // the ghost cells
  if (bc == 0) { g0 = Cn; gL = Cf; }
  else if (bc == 1) { if (Y) g0 = Cf; else gL = Cn; }
  else {
    s0 = signbit(u0) ? 1.0 : 0.0;
    s1 = signbit(-u0) ? 1.0 : 0.0;
    g0 = Cf * s0; gL = Cn * s1;
  }
  // interior face f (between cells f-1 and f): a = u_f m / d, the SMART branch coefficients;
  // facew takes face c + df from the stencil, face any face from memory
  auto pick = [&](int f, double uf, double &au, double &auu, double &ad, double &cup, double &cupup, double &cdn,
                  double cm2, double cm1, double c0, double cp1) {
    const bool pos = uf >= 0.0;
    if (pos) { cup = cm1; cupup = f >= 2 ? cm2 : g0; cdn = c0; }
    else { cup = c0; cupup = f + 1 <= n - 1 ? cp1 : gL; cdn = cm1; }
    smart_coef(cup, cupup, cdn, au, auu, ad);
    return pos;
  };
  auto facew = [&](int df, double &a, double &au, double &auu, double &ad, double &cup, double &cupup, double &cdn) {
    const double uf = Uw[df + 1];
    a = uf * id * (Mw[df + 2] - Mw[df + 1]);
    return pick(c + df, uf, au, auu, ad, cup, cupup, cdn, Cw[df + 1], Cw[df + 2], Cw[df + 3], Cw[df + 4]);
  };
  auto face = [&](int f, double &a, double &au, double &auu, double &ad, double &cup, double &cupup, double &cdn) {
    const double uf = U[f * st];
    a = uf * id * (Mv(f) - Mv(f - 1));
    return pick(f, uf, au, auu, ad, cup, cupup, cdn, Cv(max(f - 2, 0)), Cv(f - 1), Cv(f), Cv(min(f + 1, n - 1)));
  };
  auto diff = [&](int q, double &Kp, double &Km) {  // returns the cell's m / d^2
    const double Kq = Kv(q);
    Kp = 0.5 * (Kq + Kv(q + 1 < n ? q + 1 : 0));
    Km = 0.5 * (Kq + Kv(q >= 1 ? q - 1 : n - 1));
    return Mv(q) * id2;
  };
  auto diffw = [&](int dq, double &Kp, double &Km) {
    const double Kq = Kw[dq + 2];
    Kp = 0.5 * (Kq + Kw[dq + 3]);
    Km = 0.5 * (Kq + Kw[dq + 1]);
    return Mw[dq + 2] * id2;
  };
  double a, au, auu, ad, cu, cuu, cd;
  if (c - 1 >= 1 && !facew(-1, a, au, auu, ad, cu, cuu, cd)) A += auu * a;
  if (c >= 1) {
    const bool pos = facew(0, a, au, auu, ad, cu, cuu, cd);
    A += pos ? ad * a : au * a;
    AU[c * st] += smart_face(cu, cuu, cd) * id * (Mw[2] - Mw[1]);
  }
  if (c + 1 <= n - 1) {
    const bool pos = facew(1, a, au, auu, ad, cu, cuu, cd);
    A += pos ? au * a : ad * a;
  }
  if (c + 2 <= n - 1 && facew(2, a, au, auu, ad, cu, cuu, cd)) A += auu * a;
  // faces 0 and n: they reach cells 0, 1, n-2 and n-1 only
  double a0 = 0.0, aL = 0.0;
  if (c <= 1 || c >= n - 2) {
    const double m0 = Mv(0), mL = -Mv(n - 1);
    a0 = u0 * id * m0; aL = U[n * st] * id * mL;
    if (bc == 0) {
      int iu, iuu, idn;
      if (u0 >= 0.0) { iu = n - 1; iuu = n - 2; idn = 0; }
      else { iu = 0; iuu = 1; idn = n - 1; }
      if (c == iu || c == iuu || c == idn) {
        smart_coef(Cv(iu), Cv(iuu), Cv(idn), au, auu, ad);
        A += (c == iu ? au : c == iuu ? auu : ad) * (a0 + aL);
      }
      if (c == 0) {
        const double f0 = smart_face(Cv(iu), Cv(iuu), Cv(idn));
        AU[0] += f0 * id * m0;
        AU[n * st] += f0 * id * mL;
      }
    } else if (bc == 1) {  // x: face 0 carries nothing, face n the last cell; y: the mirror
      if (Y && c == 0) { AU[0] += Cf * id * m0; A += a0; }
      if (!Y && c == n - 1) { AU[n * st] += Cn * id * mL; A += aL; }
    } else if (c == 0) {   // the ghosts are the end faces
      AU[0] += g0 * id * m0;
      AU[n * st] += gL * id * mL;
    }
  }
  double Kp, Km;
  if (c >= 1) { const double m = diffw(-1, Kp, Km); A += Kp * m; }
  { const double m = diffw(0, Kp, Km); A -= (Kp + Km) * m; }
  if (c + 1 <= n - 1) { const double m = diffw(1, Kp, Km); A += Km * m; }
  if (c == 0 || c == n - 1) {  // the ghosts' adjoints, summed in the scan's order
    double ag0 = 0.0, agL = 0.0;
    const bool lo = c == 0;  // the near end's face and diffusion come from the stencil
    if (lo ? facew(1, a, au, auu, ad, cu, cuu, cd) : face(1, a, au, auu, ad, cu, cuu, cd)) ag0 += auu * a;
    if (!(lo ? face(n - 1, a, au, auu, ad, cu, cuu, cd) : facew(0, a, au, auu, ad, cu, cuu, cd))) agL += auu * a;
    if (bc == 2) { ag0 += a0; agL += aL; }
    { const double m = lo ? diffw(0, Kp, Km) : diff(0, Kp, Km); ag0 += Km * m; }
    { const double m = lo ? diff(n - 1, Kp, Km) : diffw(0, Kp, Km); agL += Kp * m; }
    if (bc == 0) A += c == n - 1 ? ag0 : agL;
    else if (bc == 1) { if (Y && c == 0) A += ag0; if (!Y && c == n - 1) A += agL; }
    else A += c == 0 ? s0 * ag0 : s1 * agL;
  }
  return A;
}

// z in gather form for one cell (x < nx) of one bin: w faces 1..nz-1 (face 0
// carries nothing, face nz-1 repeats nz-2, face 1 is central for upward flow).  A face
// adds to three distinct cells, so cell c takes the terms of faces c-1 .. c+2 and then the
// Kc diffusion terms of levels c-1, c, c+1, in the order of the level scan; cell c = k
// writes the bin's adjoint of w face k into awi (a volume per bin), cell nz-2 also face nz-1's.
__device__ __forceinline__ double drop_z_cell(double A, int i, int j, int c, int b, const double *__restrict__ C,
                                              const double *__restrict__ mu, const double *__restrict__ w,
                                              double *__restrict__ awi_all, const double *__restrict__ kc,
                                              const double *__restrict__ wsb, const DropGeo &g, const DropAdjP &P) {
  const int nz = g.nz;
  const long o = (long)j * g.ld + i, np = g.np, ob = (long)b * g.nv;
  const double ws = P.settling ? wsb[b] : 0.0;
  auto lv = [&](int q) { return (long)min(max(q, 0), nz - 1) * np + o; };  // clamped level q
  // the cell's stencil, loaded together: C at c-3 .. c+3, mu at c-2 .. c+2, w at c-1 .. c+3,
  // Kc at c-1 .. c+1; levels off the column are clamped (their values go unused)
  double Cw[7], Mw[5], Ww[5], Kw[3];
#pragma unroll
  for (int d = 0; d < 7; ++d) Cw[d] = C[ob + lv(c + d - 3)];
#pragma unroll
  for (int d = 0; d < 5; ++d) Mw[d] = mu[ob + lv(c + d - 2)];
#pragma unroll
  for (int d = 0; d < 5; ++d) Ww[d] = w[lv(c + d - 1)];
#pragma unroll
  for (int d = 0; d < 3; ++d) Kw[d] = kc[lv(c + d - 1)];
#pragma unroll
  for (int dk = -1; dk <= 2; ++dk) {  // faces k = c-1 .. c+2 within 1 .. nz-2
    const int k = c + dk;
    if (k < 1 || k > nz - 2) continue;
    const double wk = Ww[dk + 1] - ws;
    const double m = Mw[dk + 2] - Mw[dk + 1];
    double at = wk * P.idz * m;
    if (k == nz - 2) at += (Ww[dk + 2] - 0.0) * P.idz * (-Mw[dk + 2]);  // face nz-1 joins nz-2
    double au, auu, ad, face;
    int ku, kuu, kdn;
    if (k == 1 && wk >= 0.0) {
      face = 0.5 * (Cw[dk + 3] + Cw[dk + 2]);
      ku = 1; kdn = 0; kuu = -1; au = 0.5; ad = 0.5; auu = 0.0;
    } else {
      double cup, cupup, cdn;
      if (wk >= 0.0) { ku = k - 1; kuu = k - 2; kdn = k; cup = Cw[dk + 2]; cupup = Cw[dk + 1]; cdn = Cw[dk + 3]; }
      else { ku = k; kuu = k + 1; kdn = k - 1; cup = Cw[dk + 3]; cupup = Cw[dk + 4]; cdn = Cw[dk + 2]; }
      face = smart_face(cup, cupup, cdn);
      smart_coef(cup, cupup, cdn, au, auu, ad);
    }
    if (dk == 0) {
      double *awi = awi_all + ob;
      awi[(long)k * np + o] += face * P.idz * m;
      if (k == nz - 2) awi[(long)(nz - 1) * np + o] += face * P.idz * (-Mw[2]);
    }
    if (ku == c) A += au * at;
    else if (kdn == c) A += ad * at;
    else if (kuu == c) A += auu * at;
  }
  if (c >= 1 && c - 1 <= nz - 3) A += 0.5 * (Kw[1] + Kw[0]) * (Mw[1] * P.idz2);  // level c-1
  if (c <= nz - 2) {                                                            // level c
    const double m = Mw[2] * P.idz2;
    if (c <= nz - 3) A -= 0.5 * (Kw[2] + Kw[1]) * m;
    if (c >= 1) A -= 0.5 * (Kw[1] + Kw[0]) * m;
  }
  if (c + 1 <= nz - 2) A += 0.5 * (Kw[2] + Kw[1]) * (Mw[3] * P.idz2);           // level c+1
  return A;
}

// the three directions for one cell of one bin (blocks of 64 x by 4 y cells, grid z over
// level + nz bin): x, then y (levels below nz-1), then z, each as the passes above
__global__ void __launch_bounds__(256) k_adj_drop_xyz(const double *__restrict__ C, const double *__restrict__ mu,
                                                      double *__restrict__ ac, const double *__restrict__ ui0,
                                                      const double *__restrict__ vi0, const double *__restrict__ w,
                                                      double *__restrict__ aui, double *__restrict__ avi,
                                                      double *__restrict__ awi, const double *__restrict__ kc,
                                                      const double *__restrict__ wsb, DropGeo g, DropAdjP P) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int b = (int)blockIdx.z / g.nz, k = (int)blockIdx.z - b * g.nz;
  const long s = (long)b * g.nv + (long)k * g.np + (long)j * g.ld + i;
  double A = ac[s];
  if (k < g.nz - 1) {
    A = drop_line_cell<false>(A, i, j, k, b, C, mu, ui0, aui, kc, g, P);
    A = drop_line_cell<true>(A, i, j, k, b, C, mu, vi0, avi, kc, g, P);
  }
  A = drop_z_cell(A, i, j, k, b, C, mu, w, awi, kc, wsb, g, P);
  ac[s] = A;
}

// the face-velocity reconstruction transposed, along x:
// ui0(i) = mean(u) + P(i) - mean(P), P(i) = sum_{m<i} dx dudx(m), ui0(nx) = ui0(0).  With
// l the adjoint of the faces (face nx folded into 0) and S = sum l:
//   adjoint of dudx(m), m < nx-1: dx (sum_{i>m} l(i) - S (nx-1-m)/nx);  of u(m): S/nx.
// The same along y.
// The face adjoints of all bins summed (bin by bin in index order, from zero), for
// k_adj_face_line: one coalesced pass instead of npc strided loads per face inside the
// line passes.
__global__ void k_adj_face_bins(const double *__restrict__ aui, const double *__restrict__ avi,
                                double *__restrict__ lu, double *__restrict__ lv, DropGeo g) {
  const long fz = g.fp * g.nz;
  const long s = gtid();
  if (s >= fz) return;
  double a = 0.0, b = 0.0;
  for (int q = 0; q < g.npc; ++q) {
    a += aui[(long)q * fz + s];
    b += avi[(long)q * fz + s];
  }
  lu[s] = a;
  lv[s] = b;
}

// lsum: k_adj_face_bins's sums.  One line per wave (x: a (y, level) row, y: an (x, level)
// column, consecutive waves at consecutive x): S by a wave sum, the tails sum_{i>m} l(i)
// by suffix scans over 64-face chunks from the end of the line.
template <bool Y>
__global__ void __launch_bounds__(256) k_adj_face_line(const double *__restrict__ lsum, double *__restrict__ ad,
                                                       double *__restrict__ af, DropGeo g, double d) {
  const int lane = threadIdx.x & 63;
  const long line = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nlat = Y ? g.nx : g.ny;
  if (line >= (long)nlat * g.nz) return;  // uniform over the wave
  const int k = (int)(line / nlat), q = (int)(line % nlat), n = Y ? g.ny : g.nx;
  const long st = Y ? g.ld : 1, ol = Y ? (long)q : (long)q * g.ld;
  const double *L0 = lsum + (long)k * g.fp + ol;
  const long o = (long)k * g.np + ol;
  double part = 0.0;
  for (int i = 1 + lane; i < n; i += 64) part += L0[i * st];
  for (int sh = 32; sh > 0; sh >>= 1) part += __shfl_xor(part, sh, 64);
  const double S = (L0[0] + L0[(long)n * st]) + part;  // face n folded into face 0
  double carry = 0.0;
  for (int base = (n - 1) / 64 * 64; base >= 0; base -= 64) {
    const int m = base + lane;
    double v = m + 1 <= n - 1 ? L0[(long)(m + 1) * st] : 0.0;
    for (int sh = 1; sh < 64; sh <<= 1) {
      const double y = __shfl_down(v, sh, 64);
      if (lane + sh < 64) v += y;
    }
    const double tail = v + carry;  // sum_{i > m} l(i)
    carry = __shfl(tail, 0, 64);
    if (m < n - 1) ad[o + m * st] += d * (tail - S * (double)(n - 1 - m) / n);
    if (m < n) af[o + m * st] += S / n;
  }
}

// afw += the w-face adjoints summed over the bins
__global__ void k_adj_face_w(const double *__restrict__ awi, double *__restrict__ afw, DropGeo g) {
  const long s = gtid();
  if (s >= g.nv) return;
  double v = 0.0;
  for (int b = 0; b < g.npc; ++b) v += awi[(long)b * g.nv + s];
  afw[s] += v;
}

}  // namespace lp
