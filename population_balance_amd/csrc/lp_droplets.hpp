// lp_droplets.hpp -- kernels of the multi-bin droplet concentration step
// (scalars_module_crossflow_mod.f90:1494-1843, 1894-4508: pollen_all_in_one2 with
// pollen_RHS_calc2, calcbeta_pcon, step_pollen2) and of the acceleration update that
// feeds its inertia term (main.f90:864-898).
//
// The droplet state (PCon, RHS_PCon, RHS_PCon_f, Rhs_pbe) and all droplet arithmetic are
// binary64 in both precisions; T is the velocity-field type.  Bin b of a droplet field
// is the (ld, ny, nz) volume at offset b*nv.
#pragma once
#include "lp_kernels.hpp"

namespace lp {

struct DropGeo {
  int nx, ny, nz, ld, npc;
  long np;  // ld*ny
  long nv;  // ld*ny*nz
  long fp;  // ld*(ny+1): one level of a face-velocity array
};

// face velocity reconstruction, :2054-2081 (u) and :2088-2122 (v).  The reference forms
// u_int0 = matrix_x^T (dx dudx(1:nx-1), sum u); that matrix product is
//   u_int0(i) = mean(u) + P(i) - mean(P),  P(i) = sum_{m<i} dx dudx(m),
// i.e. the exclusive prefix sum of dx dudx shifted to the mean of u.  u_int0(nx+1) =
// u_int0(1).  One 64-lane wavefront per (y, z) row, four rows per block.
template <typename T>
__global__ __launch_bounds__(256) void k_face_u(const T *__restrict__ u, const T *__restrict__ dudx,
                                                double *__restrict__ ui0, DropGeo g, double dx) {
  const int lane = threadIdx.x & 63;
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= (long)g.ny * g.nz) return;
  const int k = (int)(row / g.ny), j = (int)(row % g.ny);
  const T *ur = u + (long)k * g.np + (long)j * g.ld;
  const T *er = dudx + (long)k * g.np + (long)j * g.ld;
  double *dst = ui0 + (long)k * g.fp + (long)j * g.ld;
  double su = 0.0, sp = 0.0, carry = 0.0;
  for (int m0 = 0; m0 < g.nx; m0 += 64) {
    const int m = m0 + lane;
    const double e = (m < g.nx - 1) ? dx * (double)er[m] : 0.0;
    double x = e;  // inclusive scan over the 64 lanes
    for (int off = 1; off < 64; off <<= 1) {
      double y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    const double P = carry + (x - e);
    if (m < g.nx) {
      dst[m] = P;
      sp += P;
      su += (double)ur[m];
    }
    carry += __shfl(x, 63, 64);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    su += __shfl_xor(su, off, 64);
    sp += __shfl_xor(sp, off, 64);
  }
  const double A = su / g.nx - sp / g.nx;
  for (int m = lane; m < g.nx; m += 64) dst[m] = A + dst[m];
  if (lane == 0) dst[g.nx] = A + 0.0;
}

// the same reconstruction along y; one thread per (x, z) column (coalesced over x)
template <typename T>
__global__ void k_face_v(const T *__restrict__ v, const T *__restrict__ dvdy, double *__restrict__ vi0, DropGeo g,
                         double dy) {
  const long t = gtid();
  if (t >= (long)g.nx * g.nz) return;
  const int k = (int)(t / g.nx), i = (int)(t % g.nx);
  const T *vc = v + (long)k * g.np + i;
  const T *ec = dvdy + (long)k * g.np + i;
  double *dst = vi0 + (long)k * g.fp + i;
  double su = 0.0, sp = 0.0, P = 0.0;
  for (int m = 0; m < g.ny; ++m) {
    dst[(long)m * g.ld] = P;
    sp += P;
    su += (double)vc[(long)m * g.ld];
    if (m < g.ny - 1) P += dy * (double)ec[(long)m * g.ld];
  }
  const double A = su / g.ny - sp / g.ny;
  for (int m = 0; m < g.ny; ++m) dst[(long)m * g.ld] = A + dst[(long)m * g.ld];
  dst[(long)g.ny * g.ld] = A + 0.0;
}

// :3482-3496 Kc_t = Nu_t on cell centres / Sc (zero outside the cells);
// :2135-2138 magS(:,:,nz) = magS(:,:,nz-1)
template <typename T>
__global__ void k_kc(const T *__restrict__ nut, T *__restrict__ magS, double *__restrict__ kc, DropGeo g,
                     double Sc) {
  const long s = gtid();
  if (s >= g.nv) return;
  const int k = (int)(s / g.np), x = (int)(s % g.ld);
  if (k == g.nz - 1) magS[s] = magS[s - g.np];
  if (k == g.nz - 1 || x >= g.nx) {
    kc[s] = 0.0;
    return;
  }
  const double kv = (k == 0 || k == g.nz - 2) ? (double)nut[s] : ((double)nut[s] + (double)nut[s + g.np]) / 2.0;
  kc[s] = kv / Sc;
}

// SMART face value (:2877-3475) from the upwind cell, the cell upwind of it and the
// downwind cell; a zero upwind gradient takes the upwind value
__device__ __forceinline__ double smart_face(double cup, double cupup, double cdn) {
  const double den = cup - cupup;
  if (den == 0.0) return cup;
  double r = (cdn - cup) / den;
  r = fmax(0.0, fmin(fmin(2.0 * r, 0.75 * r + 0.25), 4.0));
  return cup + r * den / 2.0;
}

__device__ __forceinline__ double sign1(double x) { return signbit(x) ? -1.0 : 1.0; }

struct PollenParams {
  double idx, idy, idz;     // 1/dx, 1/dy, 1/dz
  double idx2, idy2, idz2;  // 1/dx^2 ...
  double ghat;              // g/(u*^2/z_i): PCon_acc time-scale divisor
  double psf;               // surface flux (lbcp = 1 keeps only negative values)
  double frac, frac1;       // source shares of levels src_k and src_k+1
  int bc_x, bc_y;           // 0 periodic, 1 inlet, 2 upwind-sign ghosts
  int settling, inertia;
  int src_on, src_k0;       // source active this step; its lower level (0-based)
};

// pollen_RHS_calc2 for every bin at one cell: face velocities of the bin (:2243-2276),
// ghosts (:2478-2590), SMART faces (:2877-3475), flux divergence and SGS diffusion
// (:3559-3677) and the point source (:4222-4256).  One thread per point of the volume;
// points outside the cells keep the values of the previous RHS: when the caller has
// moved it to RF (RHS_PCon_f = RHS_PCon, :1521, done by exchanging the two buffers) they
// are copied back from RF, otherwise (RF null) they are left alone.
template <typename T>
__global__ __launch_bounds__(256) void k_pollen_rhs(const double *__restrict__ PC, double *__restrict__ RH,
                                                    const double *__restrict__ RF, const double *__restrict__ ui0,
                                                    const double *__restrict__ vi0, const T *__restrict__ w,
                                                    const T *__restrict__ dudt, const T *__restrict__ dvdt,
                                                    const T *__restrict__ dwdt, const double *__restrict__ kc,
                                                    const double *__restrict__ wsb, const double *__restrict__ qsrc,
                                                    const double *__restrict__ foot, DropGeo g, PollenParams P) {
  const long s = gtid();
  if (s >= g.nv) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long np = g.np, ld = g.ld;
  const int k = (int)(s / np);
  const int j = (int)((s % np) / ld);
  const int i = (int)(s % ld);
  if (k >= nz - 1 || i >= nx) {
    if (RF)
      for (int b = 0; b < g.npc; ++b) RH[(long)b * g.nv + s] = RF[(long)b * g.nv + s];
    return;
  }
  const int im = (i == 0) ? nx - 1 : i - 1, ip = (i == nx - 1) ? 0 : i + 1;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const long row = (long)k * np + (long)j * ld;  // start of this x row
  const long col = (long)k * np + i;             // start of this y column
  // bin-independent face velocities: faces i, i+1 and 1 of the row / column
  const double *ur = ui0 + (long)k * g.fp + (long)j * ld;
  const double *vc = vi0 + (long)k * g.fp + i;
  const double u0 = ur[i], u1 = ur[i + 1], uF = ur[0];
  const double v0 = vc[(long)j * ld], v1 = vc[(long)(j + 1) * ld], vF = vc[0];
  const double w0 = (double)w[s], w1 = (double)w[s + np];
  double ax0 = 0.0, ax1 = 0.0, axF = 0.0, ay0 = 0.0, ay1 = 0.0, ayF = 0.0, az0 = 0.0, az1 = 0.0;
  const bool inert = P.settling && P.inertia;
  if (inert) {  // accelerations on the faces (:2255-2270)
    const T *a = dudt + row;
    axF = 0.5 * ((double)a[nx - 1] + (double)a[0]);
    ax0 = (i == 0) ? axF : 0.5 * ((double)a[i - 1] + (double)a[i]);
    ax1 = (i == nx - 1) ? axF : 0.5 * ((double)a[i] + (double)a[i + 1]);
    const T *bcol = dvdt + col;
    ayF = 0.5 * ((double)bcol[(long)(ny - 1) * ld] + (double)bcol[0]);
    ay0 = (j == 0) ? ayF : 0.5 * ((double)bcol[(long)(j - 1) * ld] + (double)bcol[(long)j * ld]);
    ay1 = (j == ny - 1) ? ayF : 0.5 * ((double)bcol[(long)j * ld] + (double)bcol[(long)(j + 1) * ld]);
    az0 = (double)dwdt[s];
    az1 = (double)dwdt[s + np];
  }
  // Kc_t stencil (same for every bin)
  const double K0 = kc[s];
  const double Kxm = kc[row + im], Kxp = kc[row + ip];
  const double Kym = kc[col + (long)jm * ld], Kyp = kc[col + (long)jp * ld];
  const double Kzm = (k > 0) ? kc[s - np] : 0.0;
  const double Kzp = (k < nz - 2) ? kc[s + np] : 0.0;
  double fsrc = 0.0;
  bool src_here = false;
  if (P.src_on && (k == P.src_k0 || k == P.src_k0 + 1)) {
    src_here = true;
    fsrc = foot[(long)j * nx + i];
  }
  const double fr = (k == P.src_k0) ? P.frac : P.frac1;

  for (int b = 0; b < g.npc; ++b) {
    const double *C = PC + (long)b * g.nv;
    const double ws = wsb[b];
    double uf0 = u0, uf1 = u1, ufF = uF, vf0 = v0, vf1 = v1, vfF = vF, wf0 = w0, wf1 = w1;
    if (P.settling) {  // w faces 2..nz-1 of the reference carry the settling velocity
      if (k >= 1) wf0 -= ws;
      if (k + 1 <= nz - 2) wf1 -= ws;
      if (P.inertia) {
        const double f = ws / P.ghat;
        uf0 -= ax0 * f; uf1 -= ax1 * f; ufF -= axF * f;
        vf0 -= ay0 * f; vf1 -= ay1 * f; vfF -= ayF * f;
        if (k >= 1) wf0 -= az0 * f;
        if (k + 1 <= nz - 2) wf1 -= az1 * f;
      }
    }
    auto Cx = [&](int ii) -> double { return C[row + ii]; };
    auto Cy = [&](int jj) -> double { return C[col + (long)jj * ld]; };
    auto Cz = [&](int kk) -> double { return C[(long)kk * np + (long)j * ld + i]; };
    const double ci = C[s];
    // ghost cells -1 and n of the row / column
    double gx0, gxL, gy0, gyL;
    if (P.bc_x == 0) { gx0 = Cx(nx - 1); gxL = Cx(0); }
    else if (P.bc_x == 1) { gx0 = 0.0; gxL = Cx(nx - 1); }
    else { gx0 = Cx(0) * (1 - sign1(ufF)) / 2; gxL = Cx(nx - 1) * (1 - sign1(-ufF)) / 2; }
    if (P.bc_y == 0) { gy0 = Cy(ny - 1); gyL = Cy(0); }
    else if (P.bc_y == 1) { gy0 = Cy(0); gyL = 0.0; }
    else { gy0 = Cy(0) * (1 - sign1(vfF)) / 2; gyL = Cy(ny - 1) * (1 - sign1(-vfF)) / 2; }
    // face f lies between cells f-1 and f; end faces of a periodic row use the face-1 sign
    auto facex = [&](int f, double uf) -> double {
      if (f == 0 || f == nx) {
        if (P.bc_x == 0)
          return (ufF >= 0.0) ? smart_face(Cx(nx - 1), Cx(nx - 2), Cx(0)) : smart_face(Cx(0), Cx(1), Cx(nx - 1));
        return (f == 0) ? gx0 : gxL;
      }
      const double cm2 = (f >= 2) ? Cx(f - 2) : gx0;
      const double cp1 = (f + 1 <= nx - 1) ? Cx(f + 1) : gxL;
      return (uf >= 0.0) ? smart_face(Cx(f - 1), cm2, Cx(f)) : smart_face(Cx(f), cp1, Cx(f - 1));
    };
    auto facey = [&](int f, double vf) -> double {
      if (f == 0 || f == ny) {
        if (P.bc_y == 0)
          return (vfF >= 0.0) ? smart_face(Cy(ny - 1), Cy(ny - 2), Cy(0)) : smart_face(Cy(0), Cy(1), Cy(ny - 1));
        return (f == 0) ? gy0 : gyL;
      }
      const double cm2 = (f >= 2) ? Cy(f - 2) : gy0;
      const double cp1 = (f + 1 <= ny - 1) ? Cy(f + 1) : gyL;
      return (vf >= 0.0) ? smart_face(Cy(f - 1), cm2, Cy(f)) : smart_face(Cy(f), cp1, Cy(f - 1));
    };
    // z faces 1..nz-2 (0-based); face 0 carries no flux, face 1 is central for upward
    // flow, and the top face nz-1 repeats face nz-2 (:3340, :3356)
    auto facez = [&](int f, double wf) -> double {
      if (f == 0) return 0.0;
      if (f == 1) return (wf >= 0.0) ? (Cz(1) + Cz(0)) / 2.0 : smart_face(Cz(1), Cz(2), Cz(0));
      return (wf >= 0.0) ? smart_face(Cz(f - 1), Cz(f - 2), Cz(f)) : smart_face(Cz(f), Cz(f + 1), Cz(f - 1));
    };
    const double cxb = facex(i, uf0), cxt = facex(i + 1, uf1);
    const double cyb = facey(j, vf0), cyt = facey(j + 1, vf1);
    const double czb = facez(k, wf0);
    const double czt = (k + 1 == nz - 1) ? czb : facez(k + 1, wf1);
    double r = -P.idx * (uf1 * cxt - uf0 * cxb);
    r = r - P.idy * (vf1 * cyt - vf0 * cyb);
    r = r - P.idz * (wf1 * czt - wf0 * czb);
    {
      const double cm = (i == 0) ? gx0 : Cx(i - 1);
      const double cp = (i == nx - 1) ? gxL : Cx(i + 1);
      r = r + P.idx2 * (((K0 + Kxp) / 2.0) * (cp - ci) - ((K0 + Kxm) / 2.0) * (ci - cm));
    }
    {
      const double cm = (j == 0) ? gy0 : Cy(j - 1);
      const double cp = (j == ny - 1) ? gyL : Cy(j + 1);
      r = r + P.idy2 * (((K0 + Kyp) / 2.0) * (cp - ci) - ((K0 + Kym) / 2.0) * (ci - cm));
    }
    if (k == 0)
      r = r + P.idz2 * (((Kzp + K0) / 2.0) * (Cz(1) - ci)) + P.idz * P.psf;
    else if (k == nz - 2)
      r = r + P.idz2 * (0.0 - ((K0 + Kzm) / 2.0) * (ci - Cz(k - 1)));
    else
      r = r + P.idz2 * (((Kzp + K0) / 2.0) * (Cz(k + 1) - ci) - ((K0 + Kzm) / 2.0) * (ci - Cz(k - 1)));
    if (src_here) r += fr * qsrc[b] * fsrc;
    RH[(long)b * g.nv + s] = r;
  }
}

// breakage-frequency fit of one bin (:2167-2230): the Ohnesorge-number coefficients and
// the two diameter powers
struct PbeBin {
  double a, b, c, d, e;
  double dpow_tau;  // d^-0.6666667
  double dpow_re;   // d^(4/3)
};
// one (smaller bin i, larger bin j) pair of the transfer (:4356-4448): a drop of bin j
// breaks into bin i with probability beta; the complementary drop of diameter
// (d_j^3 - d_i^3)^(1/3) is split between bins lo and hi with volume weights
struct PbePair {
  int i, j, lo, hi;  // lo/hi < 0: no target
  double beta, wlo, whi;
};

// Rhs_pbe for every bin at one cell; zero outside the cells (the reference clears the
// whole array, :2005-2021).  With feedback, RHS_PCon += Rhs_pbe.  The per-thread bin
// arrays live in LDS, laid out [bin][thread].
template <typename T>
__global__ __launch_bounds__(64) void k_pbe(const double *__restrict__ PC, double *__restrict__ PB,
                                            double *__restrict__ RH, const T *__restrict__ dissip,
                                            const PbeBin *__restrict__ bins, const PbePair *__restrict__ pairs,
                                            int npairs, DropGeo g, double kpref, double ustar3, double z_i,
                                            double nu_c, int feedback) {
  extern __shared__ double lds[];
  double *gf = lds;                     // [npc][64]
  double *acc = lds + 64 * g.npc;       // [npc][64]
  const int t = threadIdx.x;
  const long s = gtid();
  if (s >= g.nv) return;
  const int k = (int)(s / g.np), x = (int)(s % g.ld);
  if (k >= g.nz - 1 || x >= g.nx) {
    for (int b = 0; b < g.npc; ++b) PB[(long)b * g.nv + s] = 0.0;
    return;
  }
  const double eps = (double)dissip[s];
  const double ds = eps * ustar3 / z_i;
  const double et = pow(fabs(eps), 0.33333333);
  const double er = pow(fabs(ds), 1.0 / 3.0);
  for (int b = 0; b < g.npc; ++b) {
    const PbeBin B = bins[b];
    const double tau_c = et * B.dpow_tau;
    const double Re = er * B.dpow_re / nu_c;
    const double Re1 = log10(Re);
    const double fb = 0.5 * (sign1(Re1) + 1.0);
    const double al = B.a * fb * pow(Re1, fb * B.b);
    const double bl = fb * B.c * pow(Re1, fb * B.d) - B.e;
    const double fit = pow(10.0, al + bl);
    gf[b * 64 + t] = kpref * tau_c * fit;
    acc[b * 64 + t] = 0.0;
  }
  for (int q = 0; q < npairs; ++q) {
    const PbePair Q = pairs[q];
    const double tr = gf[Q.j * 64 + t] * Q.beta * PC[(long)Q.j * g.nv + s];
    acc[Q.i * 64 + t] += tr;
    acc[Q.j * 64 + t] -= tr;
    if (Q.lo >= 0) acc[Q.lo * 64 + t] += tr * Q.wlo;
    if (Q.hi >= 0) acc[Q.hi * 64 + t] += tr * Q.whi;
  }
  for (int b = 0; b < g.npc; ++b) {
    const long o = (long)b * g.nv + s;
    const double pb = acc[b * 64 + t];
    PB[o] = pb;
    if (feedback) RH[o] = RH[o] + pb;
  }
}

// calcbeta_pcon (:1795-1843; buoyancy of all bins into beta_pcon, zero outside the cells
// or when the droplets are passive) and step_pollen2 (:4481-4508; AB2 in the cells, the
// top level copies level nz-1).  One thread per point of the volume.
// do_beta / do_step select either routine or both (fused: one pass over PCon).
template <typename T>
__global__ void k_beta_step(double *__restrict__ PC, const double *__restrict__ RH, const double *__restrict__ RF,
                            T *__restrict__ bp, const double *__restrict__ vfac, int active, DropGeo g, double dt,
                            int do_beta, int do_step) {
  const long s = gtid();
  if (s >= g.nv) return;
  const int k = (int)(s / g.np), x = (int)(s % g.ld);
  if (k == g.nz - 1) {
    if (do_beta) bp[s] = T(0);
    return;
  }
  if (x >= g.nx) {
    if (do_beta) bp[s] = T(0);
    if (do_step && k == g.nz - 2)
      for (int b = 0; b < g.npc; ++b) PC[(long)b * g.nv + s + g.np] = PC[(long)b * g.nv + s];
    return;
  }
  double tb = 0.0;
  for (int b = 0; b < g.npc; ++b) {
    const long o = (long)b * g.nv + s;
    const double c = PC[o];
    if (do_beta) tb = tb + vfac[b] * c;
    if (do_step) {
      const double cn = c + dt * (1.5 * RH[o] - 0.5 * RF[o]);
      PC[o] = cn;
      if (k == g.nz - 2) PC[o + g.np] = cn;
    }
  }
  if (do_beta) bp[s] = active ? (T)tb : T(0);
}

// energy.f90:23-31 kinetic energy on the uvp nodes (zero outside the cells)
template <typename T>
__global__ void k_energy(const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
                         T *__restrict__ ke, Geo g) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (k == g.nz - 1 || x >= g.nx) {
    ke[i] = T(0);
    return;
  }
  const T tw = T(0.5) * (w[i] + w[i + g.np]);
  ke[i] = T(0.5) * (u[i] * u[i] + v[i] * v[i] + tw * tw);
}

// main.f90:883-885 with ddz_uv(dkedz, ke) folded in (ddz_uv.f90:39-57: level 1 and the
// columns beyond nx add nothing)
template <typename T>
__global__ void k_ke_add(T *__restrict__ dudt, T *__restrict__ dvdt, T *__restrict__ dwdt,
                         const T *__restrict__ kex, const T *__restrict__ key, const T *__restrict__ ke, Geo g,
                         T idz) {
  const long i = gtid();
  if (i >= g.np * (g.nz - 1)) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  dudt[i] += kex[i];
  dvdt[i] += key[i];
  const T kz = (k >= 1 && x < g.nx) ? idz * (ke[i] - ke[i - g.np]) : T(0);
  dwdt[i] += kz;
}

}  // namespace lp
