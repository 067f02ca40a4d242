// Potential temperature (S_FLAG) kernels: scalars_module_crossflow_mod.f90 obukhov
// (:1170-1403), theta_all_in_one / scalar_RHS_calc (:141-1110), calcbeta (:361-473) and
// step_scalar (:1111-1169) for a homogeneous surface with a prescribed flux.
//
// The plane reductions (obukhov's means, the patch mean of u*, calcbeta's level means)
// run in one workgroup per plane: each lane sums a fixed strided subset in binary64 and
// a fixed LDS tree combines them, so results do not depend on scheduling.  Everything
// else is pointwise over (ld, ny, nz) or the 3/2 grid, one lane per node.
#pragma once
#include "lp_kernels.hpp"

namespace lp {

// parameters of k_obukhov (host-side values of the configuration)
struct ObkP {
  int stab;        // 1: stability branch (theta on and jt >= scal_init); 0: neutral
  int zero_psi;    // neutral branch without initsc: psi_m = 0 first
  int set_t3;      // jt == scal_init without initsc: sgs_t3(:,:,1) = t3val first
  double t3val;    // wt_s / u_star / T_scale
  double vonk, dz;
  double lz0;      // log(0.5 dz / z0) (neutral branch)
  double lzo_;     // log(0.5 dz / zo_), zo_ = exp(mean(log z0)) (stability branch)
  double g_;       // g / (u_star^2 / z_i)
};

// fixed-order block sum of 1024 lanes' partials (red: 1024 doubles of LDS); lane 0
// returns the total
__device__ __forceinline__ double block_sum1024(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// :1170-1403 obukhov, one workgroup of 1024 lanes over the level-1 plane.  ob[0..7]:
// psi_m, psi_h, phi_m, phi_h, u*, L, <sgs_t3(1)>, <theta(1)>.
template <typename T>
__global__ __launch_bounds__(1024) void k_obukhov(const double *__restrict__ u, const double *__restrict__ v,
                                                  const double *__restrict__ th, double *__restrict__ t3,
                                                  double *__restrict__ psi, T *__restrict__ ust,
                                                  double *__restrict__ ob, Geo g, ObkP P) {
  __shared__ double red[1024];
  __shared__ double res[2];
  const long nxy = (long)g.nx * g.ny;
  const double dnxy = (double)nxy;
  if (P.set_t3)  // :1195-1201
    for (long i = threadIdx.x; i < g.np; i += blockDim.x) t3[i] = P.t3val;
  if (P.zero_psi)  // :1184
    for (long i = threadIdx.x; i < g.np; i += blockDim.x) psi[i] = 0.0;
  __syncthreads();
  double su = 0.0, sp = 0.0, st = 0.0, sh = 0.0;
  for (long q = threadIdx.x; q < nxy; q += blockDim.x) {
    const long i = (q / g.nx) * g.ld + q % g.nx;
    su += sqrt(u[i] * u[i] + v[i] * v[i]);
    sp += psi[i];
    st += t3[i];
    sh += th ? th[i] : 0.0;
  }
  su = block_sum1024(su, red);
  sp = block_sum1024(sp, red);
  st = block_sum1024(st, red);
  sh = block_sum1024(sh, red);
  if (threadIdx.x == 0) {
    if (!P.stab) {  // :1183-1193
      const double us = su / dnxy * P.vonk / (P.lz0 - sp / dnxy);
      ob[0] = sp / dnxy; ob[1] = 0.0; ob[2] = 1.0; ob[3] = 1.0;
      ob[4] = us; ob[5] = 0.0; ob[6] = 0.0; ob[7] = 0.0;
      res[0] = 0.0;
      res[1] = us;
    } else {  // :1202-1272, homogeneous
      const double wt = st / dnxy;
      const double us = su / dnxy * P.vonk / (P.lzo_ - sp / dnxy);
      const double tm = sh / dnxy;
      const double L = -(us * us * us) / (P.vonk * P.g_ / tm * wt);
      double pm = 0.0, ph = 0.0, fm = 1.0, fh = 1.0;
      if (L < 0.0 && wt != 0.0) {
        const double x = pow(1.0 - 16.0 * P.dz / 2.0 / L, 0.25);
        pm = 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + M_PI / 2.0;
        ph = 2.0 * log((1.0 + x * x) / 2.0);
        fm = 1.0 / x;
        fh = 1.0 / (x * x);
      } else if (L > 0.0 && wt != 0.0) {
        const double zeta = 0.5 * P.dz / L;
        fm = 1.0 + 4.8 * zeta;
        fh = 1.0 + 7.8 * zeta;
        pm = -4.8 * zeta;
        ph = -7.8 * zeta;
      }
      ob[0] = pm; ob[1] = ph; ob[2] = fm; ob[3] = fh; ob[4] = us; ob[5] = L; ob[6] = wt; ob[7] = tm;
      res[0] = pm;
      res[1] = us;
    }
  }
  __syncthreads();
  for (long q = threadIdx.x; q < nxy; q += blockDim.x) {
    const long i = (q / g.nx) * g.ld + q % g.nx;
    if (P.stab) psi[i] = res[0];
    ust[i] = (T)res[1];
  }
}

// scalar_RHS_calc :530-575 (patch_flag = 1, one patch type): u* = sum of u*/(nx ny) over
// the plane, then dsdz(:,:,1) = -phi_h surf_flux / (u* vonk dz / 2); one workgroup
template <typename T>
__global__ __launch_bounds__(1024) void k_theta_sfc(const T *__restrict__ ust, const double *__restrict__ ob,
                                                    T *__restrict__ dTdz, Geo g, double sflux, double vonk,
                                                    double dz) {
  __shared__ double red[1024];
  __shared__ double val;
  const long nxy = (long)g.nx * g.ny;
  const double dnxy = (double)nxy;
  double s = 0.0;
  for (long q = threadIdx.x; q < nxy; q += blockDim.x) s += (double)ust[(q / g.nx) * g.ld + q % g.nx] / dnxy;
  s = block_sum1024(s, red);
  if (threadIdx.x == 0) val = -ob[3] * sflux / (s * vonk * dz / 2.0);
  __syncthreads();
  for (long q = threadIdx.x; q < nxy; q += blockDim.x) dTdz[(q / g.nx) * g.ld + q % g.nx] = (T)val;
}

// :827-874 advection on the 3/2 grid, written over u_m (each node reads only its own
// u_m value): u_m dsdx_m + v_m dsdy_m + the w-node average of w_m dsdz_m (level 1: the
// upper node's half, level nz: none)
template <typename T>
__global__ void k_theta_adv(T *__restrict__ um, const T *__restrict__ vm, const T *__restrict__ wm,
                            const T *__restrict__ sx, const T *__restrict__ sy, const T *__restrict__ sz, int nx2,
                            int ny2, int ldb, int nz) {
  const long npb = (long)ldb * ny2;
  const long i = gtid();
  if (i >= npb * nz) return;
  const int k = (int)(i / npb), x = (int)(i % ldb);
  if (x >= nx2) return;
  T r = um[i] * sx[i] + vm[i] * sy[i];
  if (k == 0)
    r = r + (T(0.5) * wm[i + npb]) * sz[i + npb];
  else if (k < nz - 1)
    r = r + (wm[i] * sz[i] + wm[i + npb] * sz[i + npb]) / T(2);
  um[i] = r;
}

// :889-975 the x and y SGS fluxes (1/Pr) Nu_t dsdx (level 1), (1/Pr) Nu_t_uv dsdx above;
// level nz and the padding columns zero
template <typename T>
__global__ void k_theta_flux_xy(const T *__restrict__ nut, const T *__restrict__ sx, const T *__restrict__ sy,
                                T *__restrict__ fx, T *__restrict__ fy, Geo g, double inv) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  double a = 0.0, b = 0.0;
  if (x < g.nx && k < g.nz - 1) {
    const double n0 = nut[i];
    if (k == 0) {
      a = inv * n0 * (double)sx[i];
      b = inv * n0 * (double)sy[i];
    } else {
      const double nn = inv * 0.5 * (n0 + (double)nut[i + g.np]);
      a = nn * (double)sx[i];
      b = nn * (double)sy[i];
    }
  }
  fx[i] = (T)a;
  fy[i] = (T)b;
}

// :990-1020 the z flux on w nodes: -surf_flux on level 1, (1/Pr) Nu_t dsdz above; sgs_vert
// (sgs_t3) is minus it (surf_flux on level 1)
template <typename T>
__global__ void k_theta_flux_z(const T *__restrict__ nut, const T *__restrict__ sz, T *__restrict__ fz,
                               double *__restrict__ t3, Geo g, double inv, double sflux) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (x >= g.nx) {
    fz[i] = T(0);
    return;
  }
  if (k == 0) {
    fz[i] = (T)(-1.0 * sflux);
    t3[i] = sflux;
  } else {
    const double t = inv * (double)nut[i] * (double)sz[i];
    fz[i] = (T)t;
    t3[i] = -1.0 * (double)(T)t;
  }
}

// :930-1060 RHS = -advection + d/dx flux_x (levels 1..nz-1) + d/dy flux_y (all levels) +
// d/dz flux_z (levels 1..nz-1), in the reference's order of additions
template <typename T>
__global__ void k_theta_rhs(T *__restrict__ rhs, const T *__restrict__ dx, const T *__restrict__ dy,
                            const T *__restrict__ dz, Geo g) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  if (x >= g.nx) return;
  double r = rhs[i];
  if (k < g.nz - 1) r = -1.0 * r + (double)dx[i];
  r = r + (double)dy[i];
  if (k < g.nz - 1) r = r + (double)dz[i];
  rhs[i] = (T)r;
}

// calcbeta :375-383: the plane mean of each level (one workgroup per level)
static __global__ __launch_bounds__(1024) void k_level_mean(const double *__restrict__ th, Geo g,
                                                            double *__restrict__ bar) {
  __shared__ double red[1024];
  const int k = blockIdx.x;
  const long nxy = (long)g.nx * g.ny;
  double s = 0.0;
  for (long q = threadIdx.x; q < nxy; q += blockDim.x) s += th[(long)k * g.np + (q / g.nx) * g.ld + q % g.nx];
  s = block_sum1024(s, red);
  if (threadIdx.x == 0) bar[k] = s / (double)nxy;
}

// calcbeta :385-435 on w nodes 2..nz-1 (zero elsewhere): ocean (iboussinesq 1, 2, 3) or
// atmospheric buoyancy of theta
template <typename T>
__global__ void k_calcbeta(const double *__restrict__ th, const double *__restrict__ bar, T *__restrict__ beta,
                           Geo g, int mode, double g_hat, double at) {
  const long i = gtid();
  if (i >= g.np * g.nz) return;
  const int k = (int)(i / g.np), x = (int)(i % g.ld);
  double r = 0.0;
  if (x < g.nx && k >= 1 && k < g.nz - 1) {
    const double a = th[i], b = th[i - g.np];
    double above, below;
    if (mode == 1) {
      above = 1.0 - at * a;
      below = 1.0 - at * b;
      r = -g_hat * (above + below) / 2.0;
    } else if (mode == 2) {
      above = at * (a - bar[k]) / (1.0 - at * bar[k]);
      below = at * (b - bar[k - 1]) / (1.0 - at * bar[k - 1]);
      r = g_hat * (above + below) / 2.0;
    } else if (mode == 3) {
      above = at * a;
      below = at * b;
      r = g_hat * (above + below) / 2.0;
    } else {
      above = (a - bar[k]) / bar[k];
      below = (b - bar[k - 1]) / bar[k - 1];
      r = g_hat * (above + below) / 2.0;
    }
  }
  beta[i] = (T)r;
}

// step_scalar :1111-1152: AB2 on levels 1..nz-1
template <typename T>
__global__ void k_step_scalar(double *__restrict__ th, const T *__restrict__ rhs, const T *__restrict__ rhsf, Geo g,
                              double dt) {
  const long i = gtid();
  if (i >= g.np * (g.nz - 1)) return;
  if ((int)(i % g.ld) >= g.nx) return;
  th[i] = th[i] + dt * (1.5 * (double)rhs[i] - 0.5 * (double)rhsf[i]);
}

// step_scalar :1155-1166: the top level from the one below (the whole plane)
static __global__ void k_scalar_top(double *__restrict__ th, Geo g, double add, int with_gradient) {
  const long i = gtid();
  if (i >= g.np) return;
  const long it = (long)(g.nz - 1) * g.np + i;
  th[it] = with_gradient ? th[it - g.np] + add : th[it - g.np];
}

}  // namespace lp
