// lp_droplets.hpp -- kernels of the multi-bin droplet concentration step
// (scalars_module_crossflow_mod.f90:1494-1843, 1894-4508: pollen_all_in_one2 with
// pollen_RHS_calc2, calcbeta_pcon, step_pollen2) and of the acceleration update that
// feeds its inertia term (main.f90:864-898).
//
// The droplet state (PCon, RHS_PCon, RHS_PCon_f, Rhs_pbe) and all droplet arithmetic are
// binary64 in both precisions, and so are u, v, w; T is the type of the other fields.  Bin b of a droplet field
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
__global__ __launch_bounds__(256) void k_face_u(const double *__restrict__ u, const T *__restrict__ dudx,
                                                double *__restrict__ ui0, DropGeo g, double dx) {
  const int lane = threadIdx.x & 63;
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= (long)g.ny * g.nz) return;
  const int k = (int)(row / g.ny), j = (int)(row % g.ny);
  const double *ur = u + (long)k * g.np + (long)j * g.ld;
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

// the same reconstruction along y.  A block takes 64 x-columns of one level (coalesced
// over x); its kFaceSeg wavefronts scan consecutive y segments, and the segment sums
// are combined through LDS.
static constexpr int kFaceSeg = 8;
template <typename T>
__global__ __launch_bounds__(64 * kFaceSeg) void k_face_v(const double *__restrict__ v, const T *__restrict__ dvdy,
                                                         double *__restrict__ vi0, DropGeo g, double dy) {
  __shared__ double E[kFaceSeg][64], SU[kFaceSeg][64], SP[kFaceSeg][64];
  const int x = threadIdx.x & 63, seg = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + x, k = blockIdx.y;
  const bool on = i < g.nx;
  const int L = (g.ny + kFaceSeg - 1) / kFaceSeg;
  const int m0 = seg * L, m1 = m0 + L < g.ny ? m0 + L : g.ny;
  const double *vc = v + (long)k * g.np + (on ? i : 0);
  const T *ec = dvdy + (long)k * g.np + (on ? i : 0);
  auto e_at = [&](int m) -> double { return (m < g.ny - 1) ? dy * (double)ec[(long)m * g.ld] : 0.0; };
  double P = 0.0, su = 0.0, sp = 0.0;  // segment-local exclusive prefix and sums
  for (int m = m0; m < m1; ++m) {
    sp += P;
    su += (double)vc[(long)m * g.ld];
    P += e_at(m);
  }
  E[seg][x] = P;
  SU[seg][x] = su;
  SP[seg][x] = sp;
  __syncthreads();
  double off = 0.0, tsu = 0.0, tsp = 0.0, run = 0.0;
  for (int q = 0; q < kFaceSeg; ++q) {
    const int n0 = q * L, n1 = n0 + L < g.ny ? n0 + L : g.ny;
    const int n = n1 > n0 ? n1 - n0 : 0;
    if (q == seg) off = run;
    tsu += SU[q][x];
    tsp += SP[q][x] + (double)n * run;
    run += E[q][x];
  }
  const double A = tsu / g.ny - tsp / g.ny;
  if (!on) return;
  double *dst = vi0 + (long)k * g.fp + i;
  P = off;
  for (int m = m0; m < m1; ++m) {
    dst[(long)m * g.ld] = A + P;
    P += e_at(m);
  }
  if (seg == 0) dst[(long)g.ny * g.ld] = A + 0.0;
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
  // cup + psi(r) den / 2 with r = num / den, psi(r) = max(0, min(2r, 0.75 r + 0.25, 4)),
  // without the division: psi(r) den is max(0, min(2 num, 0.75 num + 0.25 den, 4 den))
  // for den > 0, and min(0, max(...)) of the same three for den < 0
  const double den = cup - cupup;
  if (den == 0.0) return cup;
  const double num = cdn - cup;
  const double a = 2.0 * num, b = 0.75 * num + 0.25 * den, c = 4.0 * den;
  const double pd = den > 0.0 ? fmax(0.0, fmin(fmin(a, b), c)) : fmin(0.0, fmax(fmax(a, b), c));
  return cup + pd / 2.0;
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

// pollen_RHS_calc2 at one cell for a group of G bins: face velocities of the bin
// (:2243-2276), ghosts (:2478-2590), SMART faces (:2877-3475), flux divergence and SGS
// diffusion (:3559-3677) and the point source (:4222-4256).  A thread takes one point and
// G consecutive bins (blockIdx.y = bin group), so the bin-independent inputs (face
// velocities, accelerations, Kc_t) are loaded once per G bins.  Every load of a bin is
// issued unconditionally from a clamped index and the stencil logic selects among the
// loaded values; the row and column end values are loaded only where a face or ghost of
// the row or column end is formed.  Points outside the cells keep the values of the
// previous RHS: when the caller has moved it to RF (RHS_PCon_f = RHS_PCon, :1521, done by
// exchanging the two buffers) they are copied back from RF, otherwise (RF null) they are
// left alone.
static constexpr int kPollenBins = 8;  // bins per thread
template <typename T>
__global__ __launch_bounds__(256) void k_pollen_rhs(
    const double *__restrict__ PC, double *__restrict__ RH, const double *__restrict__ RF,
    const double *__restrict__ ui0, const double *__restrict__ vi0, const double *__restrict__ w,
    const T *__restrict__ dudt, const T *__restrict__ dvdt, const T *__restrict__ dwdt, const double *__restrict__ kc,
    const double *__restrict__ wsb, const double *__restrict__ qsrc, const double *__restrict__ foot, DropGeo g,
    PollenParams P) {
  const long s = gtid();
  const int b0 = blockIdx.y * kPollenBins;  // first bin of the group (uniform)
  if (s >= g.nv) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long np = g.np, ld = g.ld;
  const int k = (int)(s / np);
  const int j = (int)((s % np) / ld);
  const int i = (int)(s % ld);
  if (k >= nz - 1 || i >= nx) {
    if (RF)
#pragma unroll
      for (int bb = 0; bb < kPollenBins; ++bb)
        if (b0 + bb < g.npc) RH[(long)(b0 + bb) * g.nv + s] = RF[(long)(b0 + bb) * g.nv + s];
    return;
  }
  const int im = (i == 0) ? nx - 1 : i - 1, ip = (i == nx - 1) ? 0 : i + 1;
  const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
  const long row = (long)k * np + (long)j * ld;  // start of this x row
  const long col = (long)k * np + i;             // start of this y column
  // clamped neighbour indices (values outside the row / column are selected away below)
  const int i_m2 = i >= 2 ? i - 2 : 0, i_m1 = i >= 1 ? i - 1 : 0;
  const int i_p1 = i + 1 <= nx - 1 ? i + 1 : nx - 1, i_p2 = i + 2 <= nx - 1 ? i + 2 : nx - 1;
  const int j_m2 = j >= 2 ? j - 2 : 0, j_m1 = j >= 1 ? j - 1 : 0;
  const int j_p1 = j + 1 <= ny - 1 ? j + 1 : ny - 1, j_p2 = j + 2 <= ny - 1 ? j + 2 : ny - 1;
  const int k_m2 = k >= 2 ? k - 2 : 0, k_m1 = k >= 1 ? k - 1 : 0;
  const int k_p2 = k + 2 <= nz - 1 ? k + 2 : nz - 1;  // k + 1 <= nz - 1 always
  const long pl = (long)j * ld + i;                   // offset of this point in a level
  const bool xends = i <= 1 || i >= nx - 2;           // faces or ghosts of the row ends used
  const bool yends = j <= 1 || j >= ny - 2;
  // bin-independent inputs: faces i, i+1 and 1 of the row / column
  const double *ur = ui0 + (long)k * g.fp + (long)j * ld;
  const double *vc = vi0 + (long)k * g.fp + i;
  const double u0 = ur[i], u1 = ur[i + 1], uF = ur[0];
  const double v0 = vc[(long)j * ld], v1 = vc[(long)(j + 1) * ld], vF = vc[0];
  const double w0 = (double)w[s], w1 = (double)w[s + np];
  double ax0 = 0.0, ax1 = 0.0, axF = 0.0, ay0 = 0.0, ay1 = 0.0, ayF = 0.0, az0 = 0.0, az1 = 0.0;
  const bool inert = P.settling && P.inertia;
  if (inert) {  // accelerations on the faces (:2255-2270)
    const T *a = dudt + row;
    const double am = (double)a[i_m1], a0 = (double)a[i], ap = (double)a[i_p1];
    const double aL = (double)a[nx - 1], aR = (double)a[0];
    axF = 0.5 * (aL + aR);
    ax0 = (i == 0) ? axF : 0.5 * (am + a0);
    ax1 = (i == nx - 1) ? axF : 0.5 * (a0 + ap);
    const T *bcol = dvdt + col;
    const double bm = (double)bcol[(long)j_m1 * ld], bc0 = (double)bcol[(long)j * ld];
    const double bp = (double)bcol[(long)j_p1 * ld];
    const double bL = (double)bcol[(long)(ny - 1) * ld], bR = (double)bcol[0];
    ayF = 0.5 * (bL + bR);
    ay0 = (j == 0) ? ayF : 0.5 * (bm + bc0);
    ay1 = (j == ny - 1) ? ayF : 0.5 * (bc0 + bp);
    az0 = (double)dwdt[s];
    az1 = (double)dwdt[s + np];
  }
  // Kc_t stencil
  const double K0 = kc[s];
  const double Kxm = kc[row + im], Kxp = kc[row + ip];
  const double Kym = kc[col + (long)jm * ld], Kyp = kc[col + (long)jp * ld];
  const double kzm_v = kc[s - (k > 0 ? np : 0)], kzp_v = kc[s + np];
  const double Kzm = (k > 0) ? kzm_v : 0.0;
  const double Kzp = (k < nz - 2) ? kzp_v : 0.0;
  const bool src_here = P.src_on && (k == P.src_k0 || k == P.src_k0 + 1);
  const double fsrc = src_here ? foot[(long)j * nx + i] : 0.0;
  const double fr = (k == P.src_k0) ? P.frac : P.frac1;

#pragma unroll
  for (int bb = 0; bb < kPollenBins; ++bb) {
    const int b = b0 + bb;
    if (b >= g.npc) break;
    // this bin's concentrations
    const double *C = PC + (long)b * g.nv;
    const double ci = C[s];
    const double cx_m2 = C[row + i_m2], cx_m1 = C[row + i_m1], cx_p1 = C[row + i_p1], cx_p2 = C[row + i_p2];
    const double cy_m2 = C[col + (long)j_m2 * ld], cy_m1 = C[col + (long)j_m1 * ld];
    const double cy_p1 = C[col + (long)j_p1 * ld], cy_p2 = C[col + (long)j_p2 * ld];
    const double cz_m2 = C[(long)k_m2 * np + pl], cz_m1 = C[(long)k_m1 * np + pl];
    const double cz_p1 = C[(long)(k + 1) * np + pl], cz_p2 = C[(long)k_p2 * np + pl];
    double cx_0 = 0.0, cx_1 = 0.0, cx_n2 = 0.0, cx_n1 = 0.0, cy_0 = 0.0, cy_1 = 0.0, cy_n2 = 0.0, cy_n1 = 0.0;
    if (xends) { cx_0 = C[row]; cx_1 = C[row + 1]; cx_n2 = C[row + nx - 2]; cx_n1 = C[row + nx - 1]; }
    if (yends) {
      cy_0 = C[col]; cy_1 = C[col + ld];
      cy_n2 = C[col + (long)(ny - 2) * ld]; cy_n1 = C[col + (long)(ny - 1) * ld];
    }
    // face velocities of this bin
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
    // ghost cells -1 and n of the row / column (used only where xends / yends)
    double gx0, gxL, gy0, gyL;
    if (P.bc_x == 0) { gx0 = cx_n1; gxL = cx_0; }
    else if (P.bc_x == 1) { gx0 = 0.0; gxL = cx_n1; }
    else { gx0 = cx_0 * (1 - sign1(ufF)) / 2; gxL = cx_n1 * (1 - sign1(-ufF)) / 2; }
    if (P.bc_y == 0) { gy0 = cy_n1; gyL = cy_0; }
    else if (P.bc_y == 1) { gy0 = cy_0; gyL = 0.0; }
    else { gy0 = cy_0 * (1 - sign1(vfF)) / 2; gyL = cy_n1 * (1 - sign1(-vfF)) / 2; }
    // face f lies between cells f-1 and f; the end faces of a periodic row use the face-1
    // sign
    double cxb, cxt, cyb, cyt;
    if (i == 0) {
      cxb = (P.bc_x == 0) ? ((ufF >= 0.0) ? smart_face(cx_n1, cx_n2, cx_0) : smart_face(cx_0, cx_1, cx_n1)) : gx0;
    } else {
      const double cm2 = (i >= 2) ? cx_m2 : gx0, cp1 = (i + 1 <= nx - 1) ? cx_p1 : gxL;
      cxb = (uf0 >= 0.0) ? smart_face(cx_m1, cm2, ci) : smart_face(ci, cp1, cx_m1);
    }
    if (i + 1 == nx) {
      cxt = (P.bc_x == 0) ? ((ufF >= 0.0) ? smart_face(cx_n1, cx_n2, cx_0) : smart_face(cx_0, cx_1, cx_n1)) : gxL;
    } else {
      const double cm2 = (i + 1 >= 2) ? cx_m1 : gx0, cp1 = (i + 2 <= nx - 1) ? cx_p2 : gxL;
      cxt = (uf1 >= 0.0) ? smart_face(ci, cm2, cx_p1) : smart_face(cx_p1, cp1, ci);
    }
    if (j == 0) {
      cyb = (P.bc_y == 0) ? ((vfF >= 0.0) ? smart_face(cy_n1, cy_n2, cy_0) : smart_face(cy_0, cy_1, cy_n1)) : gy0;
    } else {
      const double cm2 = (j >= 2) ? cy_m2 : gy0, cp1 = (j + 1 <= ny - 1) ? cy_p1 : gyL;
      cyb = (vf0 >= 0.0) ? smart_face(cy_m1, cm2, ci) : smart_face(ci, cp1, cy_m1);
    }
    if (j + 1 == ny) {
      cyt = (P.bc_y == 0) ? ((vfF >= 0.0) ? smart_face(cy_n1, cy_n2, cy_0) : smart_face(cy_0, cy_1, cy_n1)) : gyL;
    } else {
      const double cm2 = (j + 1 >= 2) ? cy_m1 : gy0, cp1 = (j + 2 <= ny - 1) ? cy_p2 : gyL;
      cyt = (vf1 >= 0.0) ? smart_face(ci, cm2, cy_p1) : smart_face(cy_p1, cp1, ci);
    }
    // z faces 1..nz-2 (0-based): face 0 carries no flux, face 1 is central for upward flow,
    // and the top face nz-1 repeats face nz-2 (:3340, :3356)
    double czb;
    if (k == 0) czb = 0.0;
    else if (k == 1) czb = (wf0 >= 0.0) ? (ci + cz_m1) / 2.0 : smart_face(ci, cz_p1, cz_m1);
    else czb = (wf0 >= 0.0) ? smart_face(cz_m1, cz_m2, ci) : smart_face(ci, cz_p1, cz_m1);
    double czt;
    if (k + 1 == nz - 1) czt = czb;
    else if (k == 0) czt = (wf1 >= 0.0) ? (cz_p1 + ci) / 2.0 : smart_face(cz_p1, cz_p2, ci);
    else czt = (wf1 >= 0.0) ? smart_face(ci, cz_m1, cz_p1) : smart_face(cz_p1, cz_p2, ci);
    double r = -P.idx * (uf1 * cxt - uf0 * cxb);
    r = r - P.idy * (vf1 * cyt - vf0 * cyb);
    r = r - P.idz * (wf1 * czt - wf0 * czb);
    {
      const double cm = (i == 0) ? gx0 : cx_m1;
      const double cp = (i == nx - 1) ? gxL : cx_p1;
      r = r + P.idx2 * (((K0 + Kxp) / 2.0) * (cp - ci) - ((K0 + Kxm) / 2.0) * (ci - cm));
    }
    {
      const double cm = (j == 0) ? gy0 : cy_m1;
      const double cp = (j == ny - 1) ? gyL : cy_p1;
      r = r + P.idy2 * (((K0 + Kyp) / 2.0) * (cp - ci) - ((K0 + Kym) / 2.0) * (ci - cm));
    }
    if (k == 0)
      r = r + P.idz2 * (((Kzp + K0) / 2.0) * (cz_p1 - ci)) + P.idz * P.psf;
    else if (k == nz - 2)
      r = r + P.idz2 * (0.0 - ((K0 + Kzm) / 2.0) * (ci - cz_m1));
    else
      r = r + P.idz2 * (((Kzp + K0) / 2.0) * (cz_p1 - ci) - ((K0 + Kzm) / 2.0) * (ci - cz_m1));
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
// whole array, :2005-2021).  With feedback, RHS_PCon += Rhs_pbe.  The pair loop of
// :4356-4448 is linear in g_j = gf(j) PCon(j): Rhs_pbe(b) = sum_j A(b, j) g_j with
// A(b, j) = sum over the pairs (i, j) of beta_ij (+1 into i, -1 out of j, + w_lo / w_hi into
// the complementary drop's bins), built on the host.  NB >= npc bins are held in
// registers (loops unrolled, bins >= npc skipped); A is read through the scalar cache.
// pow(Re1, b) and pow(Re1, d) share one logarithm (exp(b ln Re1)) and pow(10, x) is
// exp10(x); both agree with pow to a few ulp.
//
// STEP (pollen_all_in_one2 with the breakup source on): the same pass then does
// calcbeta_pcon and step_pollen2 at the point (as k_beta_step, from the old PCon), and on
// an Euler start first sets RHS_PCon_f = RHS_PCon (:1660-1665).
struct PbeStep {
  double *PCw;           // PCon (the same array as PC), advanced in place at this point
  double *RF;            // RHS_PCon_f
  void *bp;              // beta_pcon (field type)
  const double *vfac;    // buoyancy factor of each bin
  int active, euler;
  double dt;
};
template <typename T, int NB, bool STEP>
__global__ __launch_bounds__(256) void k_pbe(const double *PC, double *__restrict__ PB,
                                             double *__restrict__ RH, const T *__restrict__ dissip,
                                             const PbeBin *__restrict__ bins, const double *__restrict__ A,
                                             DropGeo g, double kpref, double ustar3, double z_i, double nu_c,
                                             int feedback, PbeStep St) {
  const long s = gtid();
  if (s >= g.nv) return;
  const int npc = g.npc;
  const int k = (int)(s / g.np), x = (int)(s % g.ld);
  if (k >= g.nz - 1 || x >= g.nx) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (b < npc) PB[(long)b * g.nv + s] = 0.0;
    if constexpr (STEP) {  // as k_beta_step outside the cells
      T *bp = (T *)St.bp;
      bp[s] = T(0);
      if (k == g.nz - 2)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (b < npc) St.PCw[(long)b * g.nv + s + g.np] = St.PCw[(long)b * g.nv + s];
    }
    return;
  }
  const double eps = (double)dissip[s];
  const double ds = eps * ustar3 / z_i;
  const double et = pow(fabs(eps), 0.33333333);
  const double er = pow(fabs(ds), 1.0 / 3.0);
  double gj[NB], pc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    gj[b] = 0.0;
    pc[b] = 0.0;
    if (b < npc) {
      const PbeBin B = bins[b];
      const double tau_c = et * B.dpow_tau;
      const double Re = er * B.dpow_re / nu_c;
      const double Re1 = log10(Re);
      const double fb = 0.5 * (sign1(Re1) + 1.0);
      double pb = 1.0, pd = 1.0;  // pow(Re1, fb b), pow(Re1, fb d): 1 when fb = 0
      if (fb != 0.0) {
        const double lr = log(Re1);
        pb = exp(B.b * lr);
        pd = exp(B.d * lr);
      }
      const double al = B.a * fb * pb;
      const double bl = fb * B.c * pd - B.e;
      const double fit = exp10(al + bl);
      pc[b] = PC[(long)b * g.nv + s];
      gj[b] = kpref * tau_c * fit * pc[b];
    }
  }
  double tb = 0.0;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b < npc) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < npc) acc = fma(A[b * npc + j], gj[j], acc);
      const long o = (long)b * g.nv + s;
      PB[o] = acc;
      double rh = RH[o];
      if (feedback) {
        rh = rh + acc;
        RH[o] = rh;
      }
      if constexpr (STEP) {  // calcbeta_pcon from the old PCon, then step_pollen2
        double rf;
        if (St.euler) {
          St.RF[o] = rh;
          rf = rh;
        } else {
          rf = St.RF[o];
        }
        const double c = pc[b];
        tb = tb + St.vfac[b] * c;
        const double cn = c + St.dt * (1.5 * rh - 0.5 * rf);
        St.PCw[o] = cn;
        if (k == g.nz - 2) St.PCw[o + g.np] = cn;
      }
    }
  }
  if constexpr (STEP) {
    T *bp = (T *)St.bp;
    bp[s] = St.active ? (T)tb : T(0);
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
__global__ void k_energy(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
                         T *__restrict__ ke, Geo g) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (k == g.nz - 1 || x >= g.nx) {
    ke[i] = T(0);
    return;
  }
  const double tw = 0.5 * (w[i] + w[i + g.np]);
  ke[i] = (T)(0.5 * (u[i] * u[i] + v[i] * v[i] + tw * tw));
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
