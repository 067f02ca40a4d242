/*
 * lesgo_oracle.c -- CPU binary64 restatement of the lesgo-lineage LES hot path of
 * adiaiyer/Population_Balance (velocity RHS, SGS, pressure projection).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * Each routine names the reference file:line it restates.  Reference traps kept on
 * purpose: filt_da overwrites its input; lagrange_Sdep's REAL(4) conversions (switch
 * fp32_casts); u_lag accumulates over the jt == 1 step that skips the dynamic model;
 * BOGUS (-1234567890) on the invalid top levels.  FFTs are this file's own
 * mixed-radix DFT with FFTW2's unnormalised r2c/c2r conventions (the reference
 * links FFTW 2.1.x, which is not vendored; its DFT is the textbook one).
 *
 * Parity status of each function: see the test that pins it (tests/test_oracle_*.py).
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "lesgo_oracle.h"

typedef double complex cplx;
#define BOGUS (-1234567890.0)
#define PI_D 3.1415926535897932384626433

static __thread char g_err[256];
const char *or_last_error(void) { return g_err; }
static int fail(int code, const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

int or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* 1-D complex DFT: recursive mixed radix (4, 2, 3, 5, then any prime) with a   */
/* table of exp(-2 pi i j / N).  Unnormalised; sign -1 forward, +1 backward.   */
/* ------------------------------------------------------------------------- */
typedef struct {
  int n;
  cplx *tw;
} fplan;

static void fplan_init(fplan *p, int n) {
  p->n = n;
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int j = 0; j < n; ++j) {
    double a = 2.0 * PI_D * (double)j / (double)n;
    p->tw[j] = cos(a) - I * sin(a);
  }
}
static void fplan_free(fplan *p) { free(p->tw); p->tw = NULL; }

static inline cplx twid(const fplan *P, long j, int sign) {
  cplx w = P->tw[j % P->n];
  return sign < 0 ? w : conj(w);
}

static void dft_rec(const fplan *P, int sign, cplx *out, const cplx *in, int n,
                    int stride, int s) {
  if (n == 1) { out[0] = in[0]; return; }
  int p;
  if (n % 4 == 0) p = 4;
  else if (n % 2 == 0) p = 2;
  else if (n % 3 == 0) p = 3;
  else if (n % 5 == 0) p = 5;
  else {
    p = n;
    for (int f = 7; f * f <= n; f += 2) if (n % f == 0) { p = f; break; }
  }
  int m = n / p;
  for (int r = 0; r < p; ++r) dft_rec(P, sign, out + r * m, in + r * stride, m, stride * p, s * p);
  cplx tbuf[64];
  cplx *t = p <= 64 ? tbuf : (cplx *)malloc(sizeof(cplx) * (size_t)p);
  cplx obuf[64];
  cplx *o = p <= 64 ? obuf : (cplx *)malloc(sizeof(cplx) * (size_t)p);
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      cplx y = out[r * m + k];
      t[r] = (r == 0 || k == 0) ? y : y * twid(P, (long)s * r * k, sign);
    }
    for (int q = 0; q < p; ++q) {
      cplx acc = t[0];
      for (int r = 1; r < p; ++r) acc += t[r] * twid(P, (long)s * m * ((long)r * q % p), sign);
      o[q] = acc;
    }
    for (int q = 0; q < p; ++q) out[q * m + k] = o[q];
  }
  if (t != tbuf) free(t);
  if (o != obuf) free(o);
}

static void dft(const fplan *P, int sign, cplx *out, const cplx *in) {
  dft_rec(P, sign, out, in, P->n, 1, 1);
}

void or_cfft1(int n, int sign, const double *in_c, double *out_c) {
  fplan P;
  fplan_init(&P, n);
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  memcpy(tmp, in_c, sizeof(cplx) * (size_t)n);
  dft(&P, sign, (cplx *)out_c, tmp);
  free(tmp);
  fplan_free(&P);
}

/* ------------------------------------------------------------------------- */
/* context                                                                    */
/* ------------------------------------------------------------------------- */
struct or_ctx {
  lp_config cfg;
  lp_grid g;
  fplan px, py, px2, py2;
  double *kx, *ky, *k2, *G_test, *G_test_test;
  double *f[LP_NFIELDS];
  size_t n[LP_NFIELDS];
  int jt, jt_total;
  int F_LM_MM_init, F_QN_NN_init;
  /* persistent scratch */
  double *S[6];
  double *big[7];
  /* droplet bins (copied from the configuration) */
  double *diam, *ws, *qsrc;
  /* frozen-coefficient steps (or_set_frozen): Nu_t (volume) and the wall u* (plane) of a
   * base run, used instead of the SGS model and the wall law's |U| */
  const double *frz_nut, *frz_ust;
  /* the last obukhov call: psi_m, psi_h, phi_m, phi_h, u*, L, <sgs_t3(1)>, <theta(1)>
   * (the wall law reads phi_m; the theta RHS phi_h) */
  double obuk[8];
};

static size_t plane(const or_ctx *c) { return (size_t)c->g.ld * c->g.ny; }
static size_t vol(const or_ctx *c) { return plane(c) * c->g.nz; }

or_ctx *or_create(const lp_config *cfg) {
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->cfg = *cfg;
  c->g = cfg->grid;
  const lp_grid *g = &c->g;
  fplan_init(&c->px, g->nx);
  fplan_init(&c->py, g->ny);
  fplan_init(&c->px2, g->nx2);
  fplan_init(&c->py2, g->ny2);
  size_t ph = (size_t)g->lh * g->ny;
  c->kx = (double *)calloc(ph, sizeof(double));
  c->ky = (double *)calloc(ph, sizeof(double));
  c->k2 = (double *)calloc(ph, sizeof(double));
  c->G_test = (double *)calloc(ph, sizeof(double));
  c->G_test_test = (double *)calloc(ph, sizeof(double));
  or_init_wavenumber(c, c->kx, c->ky, c->k2);
  /* main.f90:179-183 */
  or_test_filter_init(c, 2.0 * cfg->filter_size, c->G_test);
  or_test_filter_init(c, 4.0 * cfg->filter_size, c->G_test_test);
  int npc = cfg->drop.npcon > 0 ? cfg->drop.npcon : 1;
  for (int i = 0; i < LP_NFIELDS; ++i) {
    size_t n = vol(c);
    if (i == LP_P) n = plane(c) * (g->nz + 1);
    if (i == LP_USTAR || i == LP_PSI_M) n = plane(c);
    if (i == LP_PCON || i == LP_RHS_PCON || i == LP_RHS_PCON_F || i == LP_RHS_PBE)
      n = vol(c) * npc;
    c->n[i] = n;
    c->f[i] = (double *)calloc(n, sizeof(double));
  }
  c->jt = 1;
  c->jt_total = 1;
  c->obuk[2] = c->obuk[3] = 1.0; /* neutral phi_m, phi_h */
  int np_ = cfg->drop.npcon;
  if (np_ > 0) {
    c->diam = (double *)malloc(sizeof(double) * np_);
    c->ws = (double *)malloc(sizeof(double) * np_);
    c->qsrc = (double *)malloc(sizeof(double) * np_);
    for (int i = 0; i < np_; ++i) {
      c->diam[i] = cfg->drop.diameter[i];
      c->ws[i] = cfg->drop.settling_vel[i];
      c->qsrc[i] = cfg->drop.Q_src[i];
    }
    /* the configuration keeps pointers the caller may free */
    c->cfg.drop.diameter = c->diam;
    c->cfg.drop.settling_vel = c->ws;
    c->cfg.drop.Q_src = c->qsrc;
  }
  return c;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  fplan_free(&c->px); fplan_free(&c->py); fplan_free(&c->px2); fplan_free(&c->py2);
  free(c->kx); free(c->ky); free(c->k2); free(c->G_test); free(c->G_test_test);
  for (int i = 0; i < LP_NFIELDS; ++i) free(c->f[i]);
  for (int i = 0; i < 6; ++i) free(c->S[i]);
  for (int i = 0; i < 7; ++i) free(c->big[i]);
  free(c->diam); free(c->ws); free(c->qsrc);
  free(c);
}

double *or_field(or_ctx *c, int32_t field) {
  return (field >= 0 && field < LP_NFIELDS) ? c->f[field] : NULL;
}
size_t or_field_elems(or_ctx *c, int32_t field) {
  return (field >= 0 && field < LP_NFIELDS) ? c->n[field] : 0;
}
void or_set_time(or_ctx *c, int32_t jt, int32_t jt_total) { c->jt = jt; c->jt_total = jt_total; }
void or_get_time(or_ctx *c, int32_t *jt, int32_t *jt_total) { *jt = c->jt; *jt_total = c->jt_total; }

/* ------------------------------------------------------------------------- */
/* 2-D real plane transforms (FFTW2 rfftwnd_f77_one_* semantics, in-place     */
/* layout: complex(lh,ny) overlaying real(ld,ny); unnormalised).              */
/* Forward: r2c along x, then complex along y.  Inverse: complex along y,     */
/* then c2r along x using the real parts of the kx = 0 and kx = nx/2 terms.   */
/* ------------------------------------------------------------------------- */
void or_rfft2_fwd(or_ctx *c, int big, const double *in, int pitch, double *out_c) {
  const fplan *PX = big ? &c->px2 : &c->px, *PY = big ? &c->py2 : &c->py;
  int nx = PX->n, ny = PY->n, lh = nx / 2 + 1;
  cplx *out = (cplx *)out_c;
  cplx *row = (cplx *)malloc(sizeof(cplx) * (size_t)(nx > ny ? nx : ny));
  cplx *res = (cplx *)malloc(sizeof(cplx) * (size_t)(nx > ny ? nx : ny));
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) row[x] = in[(size_t)y * pitch + x];
    dft(PX, -1, res, row);
    for (int k = 0; k < lh; ++k) out[(size_t)y * lh + k] = res[k];
  }
  for (int k = 0; k < lh; ++k) {
    for (int y = 0; y < ny; ++y) row[y] = out[(size_t)y * lh + k];
    dft(PY, -1, res, row);
    for (int y = 0; y < ny; ++y) out[(size_t)y * lh + k] = res[y];
  }
  free(row);
  free(res);
}

void or_rfft2_inv(or_ctx *c, int big, const double *in_c, double *out, int pitch) {
  const fplan *PX = big ? &c->px2 : &c->px, *PY = big ? &c->py2 : &c->py;
  int nx = PX->n, ny = PY->n, lh = nx / 2 + 1;
  const cplx *in = (const cplx *)in_c;
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)lh * ny);
  int nmax = nx > ny ? nx : ny;
  cplx *row = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
  cplx *res = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
  for (int k = 0; k < lh; ++k) {
    for (int y = 0; y < ny; ++y) row[y] = in[(size_t)y * lh + k];
    dft(PY, +1, res, row);
    for (int y = 0; y < ny; ++y) tmp[(size_t)y * lh + k] = res[y];
  }
  for (int y = 0; y < ny; ++y) {
    const cplx *t = tmp + (size_t)y * lh;
    row[0] = creal(t[0]);
    row[nx / 2] = creal(t[nx / 2]);
    for (int k = 1; k < nx / 2; ++k) {
      row[k] = t[k];
      row[nx - k] = conj(t[k]);
    }
    dft(PX, +1, res, row);
    for (int x = 0; x < nx; ++x) out[(size_t)y * pitch + x] = creal(res[x]);
    for (int x = nx; x < pitch; ++x) out[(size_t)y * pitch + x] = 0.0;
  }
  free(tmp);
  free(row);
  free(res);
}

/* ------------------------------------------------------------------------- */
/* fft.f90:51-76  init_wavenumber                                             */
/* ------------------------------------------------------------------------- */
void or_init_wavenumber(or_ctx *c, double *kx, double *ky, double *k2) {
  const lp_grid *g = &c->g;
  int lh = g->lh, ny = g->ny;
  for (int jy = 0; jy < ny; ++jy)
    for (int jx = 0; jx < lh; ++jx) {
      size_t i = (size_t)jy * lh + jx;
      double vx = (jx < lh - 1) ? (double)jx : 0.0;                       /* :56-58 */
      int m = ((jy + ny / 2) % ny) - ny / 2;                              /* :60-62 */
      double vy = (double)m;
      if (jx == lh - 1) { vx = 0.0; vy = 0.0; }                           /* :65-66 */
      if (jy == ny / 2) { vx = 0.0; vy = 0.0; }                           /* :67-68 */
      kx[i] = 2.0 * PI_D / g->L_x * vx;                                   /* :71-72 */
      ky[i] = 2.0 * PI_D / g->L_y * vy;
      k2[i] = kx[i] * kx[i] + ky[i] * ky[i];                              /* :75 */
    }
}

/* test_filtermodule.f90:29-57 (ifilter = 1, sharp spectral cutoff) */
void or_test_filter_init(or_ctx *c, double alpha, double *G) {
  const lp_grid *g = &c->g;
  int lh = g->lh, ny = g->ny;
  double delta = alpha * sqrt(g->dx * g->dy);
  double kc2 = (PI_D / delta) * (PI_D / delta);
  for (int jy = 0; jy < ny; ++jy)
    for (int jx = 0; jx < lh; ++jx) {
      size_t i = (size_t)jy * lh + jx;
      double v = 1.0 / ((double)g->nx * g->ny);
      if (c->k2[i] >= kc2) v = 0.0;
      if (jx == lh - 1 || jy == ny / 2) v = 0.0;
      G[i] = v;
    }
}

/* zero the Nyquist row/column of a (lh,ny) spectrum ("oddballs") */
static void kill_oddballs(cplx *F, int lh, int ny) {
  for (int y = 0; y < ny; ++y) F[(size_t)y * lh + lh - 1] = 0.0;
  for (int x = 0; x < lh; ++x) F[(size_t)(ny / 2) * lh + x] = 0.0;
}

/* forward transform of const*f (a plane with pitch ld) into a fresh spectrum */
static cplx *fwd_scaled(or_ctx *c, const double *f, double cst) {
  const lp_grid *g = &c->g;
  size_t np = plane(c);
  double *tmp = (double *)malloc(sizeof(double) * np);
  for (int y = 0; y < g->ny; ++y)
    for (int x = 0; x < g->ld; ++x) tmp[(size_t)y * g->ld + x] = cst * f[(size_t)y * g->ld + x];
  cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)g->lh * g->ny);
  or_rfft2_fwd(c, 0, tmp, g->ld, (double *)F);
  free(tmp);
  return F;
}

/* filt_da.f90:3-36 */
void or_filt_da(or_ctx *c, double *f, double *dfdx, double *dfdy, int nlev) {
  const lp_grid *g = &c->g;
  const double cst = 1.0 / ((double)g->nx * g->ny);
  size_t np = plane(c), nh = (size_t)g->lh * g->ny;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nlev; ++k) {
    cplx *F = fwd_scaled(c, f + k * np, cst);
    kill_oddballs(F, g->lh, g->ny);
    cplx *Dx = (cplx *)malloc(sizeof(cplx) * nh), *Dy = (cplx *)malloc(sizeof(cplx) * nh);
    for (size_t i = 0; i < nh; ++i) {
      Dy[i] = I * c->ky[i] * F[i];
      Dx[i] = I * c->kx[i] * F[i];
    }
    or_rfft2_inv(c, 0, (double *)F, f + k * np, g->ld);
    or_rfft2_inv(c, 0, (double *)Dx, dfdx + k * np, g->ld);
    or_rfft2_inv(c, 0, (double *)Dy, dfdy + k * np, g->ld);
    free(F); free(Dx); free(Dy);
  }
}

/* ddx.f90:1-31, ddy.f90:1-31, ddxy.f90:1-29 */
static void ddxy_impl(or_ctx *c, double *dfdx, double *dfdy, const double *f, int nlev) {
  const lp_grid *g = &c->g;
  const double cst = 1.0 / ((double)g->nx * g->ny);
  size_t np = plane(c), nh = (size_t)g->lh * g->ny;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nlev; ++k) {
    cplx *F = fwd_scaled(c, f + k * np, cst);
    kill_oddballs(F, g->lh, g->ny);
    cplx *D = (cplx *)malloc(sizeof(cplx) * nh);
    if (dfdy) {
      for (size_t i = 0; i < nh; ++i) D[i] = I * c->ky[i] * F[i];
      or_rfft2_inv(c, 0, (double *)D, dfdy + k * np, g->ld);
    }
    if (dfdx) {
      for (size_t i = 0; i < nh; ++i) D[i] = I * c->kx[i] * F[i];
      or_rfft2_inv(c, 0, (double *)D, dfdx + k * np, g->ld);
    }
    free(F); free(D);
  }
}
void or_ddx(or_ctx *c, double *dfdx, const double *f, int nlev) { ddxy_impl(c, dfdx, NULL, f, nlev); }
void or_ddy(or_ctx *c, double *dfdy, const double *f, int nlev) { ddxy_impl(c, NULL, dfdy, f, nlev); }
void or_ddxy(or_ctx *c, double *dfdx, double *dfdy, const double *f, int nlev) {
  ddxy_impl(c, dfdx, dfdy, f, nlev);
}

/* ddz_uv.f90:5-60: dfdz(jz) = (f(jz)-f(jz-1))/dz, jz = 2..nz-1; dfdz(nz) = 0;
 * level 1 untouched */
void or_ddz_uv(or_ctx *c, double *dfdz, const double *f) {
  const lp_grid *g = &c->g;
  size_t np = plane(c);
  const double cst = 1.0 / g->dz;
#pragma omp parallel for
  for (int k = 1; k < g->nz - 1; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = k * np + (size_t)y * g->ld + x;
        dfdz[i] = cst * (f[i] - f[i - np]);
      }
  for (size_t i = 0; i < np; ++i) dfdz[(g->nz - 1) * np + i] = 0.0;
}

/* ddz_w.f90:6-51: dfdz(jz) = (f(jz+1)-f(jz))/dz, jz = 1..nz-1; dfdz(nz) = 0 */
void or_ddz_w(or_ctx *c, double *dfdz, const double *f) {
  const lp_grid *g = &c->g;
  size_t np = plane(c);
  const double cst = 1.0 / g->dz;
#pragma omp parallel for
  for (int k = 0; k < g->nz - 1; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = k * np + (size_t)y * g->ld + x;
        dfdz[i] = cst * (f[i + np] - f[i]);
      }
  for (size_t i = 0; i < np; ++i) dfdz[(g->nz - 1) * np + i] = 0.0;
}

/* padd.f90:1-26 on one complex plane */
static void padd_plane(const lp_grid *g, cplx *ub, const cplx *u) {
  int lh = g->lh, lhb = g->lh_big, nx = g->nx, ny = g->ny, ny2 = g->ny2;
  for (size_t i = 0; i < (size_t)lhb * ny2; ++i) ub[i] = 0.0;
  for (int jy = 0; jy < ny / 2; ++jy)
    for (int jx = 0; jx < nx / 2; ++jx) ub[(size_t)jy * lhb + jx] = u[(size_t)jy * lh + jx];
  for (int jy = 1; jy <= ny / 2 - 1; ++jy)
    for (int jx = 0; jx < nx / 2; ++jx)
      ub[(size_t)(jy + ny2 - ny / 2) * lhb + jx] = u[(size_t)(jy + ny / 2) * lh + jx];
}

/* unpadd.f90:1-22 on one complex plane */
static void unpadd_plane(const lp_grid *g, cplx *cc, const cplx *cb) {
  int lh = g->lh, lhb = g->lh_big, nx = g->nx, ny = g->ny, ny2 = g->ny2;
  for (int jy = 0; jy < ny / 2; ++jy)
    for (int jx = 0; jx < nx / 2; ++jx) cc[(size_t)jy * lh + jx] = cb[(size_t)jy * lhb + jx];
  for (int jy = 0; jy < ny; ++jy) cc[(size_t)jy * lh + nx / 2] = 0.0;
  for (int jx = 0; jx < lh; ++jx) cc[(size_t)(ny / 2) * lh + jx] = 0.0;
  for (int jy = 1; jy <= ny / 2 - 1; ++jy)
    for (int jx = 0; jx < nx / 2; ++jx)
      cc[(size_t)(jy + ny / 2) * lh + jx] = cb[(size_t)(jy + ny2 - ny / 2) * lhb + jx];
}

void or_padd(or_ctx *c, double *u_big, const double *u, int nplanes) {
  const lp_grid *g = &c->g;
  for (int k = 0; k < nplanes; ++k)
    padd_plane(g, (cplx *)u_big + (size_t)k * g->lh_big * g->ny2,
               (const cplx *)u + (size_t)k * g->lh * g->ny);
}
void or_unpadd(or_ctx *c, double *cc, const double *cc_big, int nplanes) {
  const lp_grid *g = &c->g;
  for (int k = 0; k < nplanes; ++k)
    unpadd_plane(g, (cplx *)cc + (size_t)k * g->lh * g->ny,
                 (const cplx *)cc_big + (size_t)k * g->lh_big * g->ny2);
}

/* dealias1.f90:1-31: u_big = c2r_big(padd(r2c(u/(nx ny)))) */
void or_dealias1(or_ctx *c, const double *u, double *u_big, int nlev) {
  const lp_grid *g = &c->g;
  const double cst = 1.0 / ((double)g->nx * g->ny);
  size_t np = plane(c), npb = (size_t)g->ld_big * g->ny2;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nlev; ++k) {
    cplx *F = fwd_scaled(c, u + k * np, cst);
    cplx *B = (cplx *)malloc(sizeof(cplx) * (size_t)g->lh_big * g->ny2);
    padd_plane(g, B, F);
    or_rfft2_inv(c, 1, (double *)B, u_big + k * npb, g->ld_big);
    free(F); free(B);
  }
}

/* dealias2.f90:1-33: u_big *= 1/(nx2 ny2); u = c2r(unpadd(r2c_big(u_big)));
 * u_big keeps its scaled spectrum (in-place transform, as in the reference). */
void or_dealias2(or_ctx *c, double *u, double *u_big, int nlev) {
  const lp_grid *g = &c->g;
  const double cst = 1.0 / ((double)g->nx2 * g->ny2);
  size_t np = plane(c), npb = (size_t)g->ld_big * g->ny2;
  for (size_t i = 0; i < npb * nlev; ++i) u_big[i] *= cst;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nlev; ++k) {
    double *ub = u_big + k * npb;
    or_rfft2_fwd(c, 1, ub, g->ld_big, ub);
    cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)g->lh * g->ny);
    unpadd_plane(g, F, (const cplx *)ub);
    or_rfft2_inv(c, 0, (double *)F, u + k * np, g->ld);
    free(F);
  }
}

/* test_filtermodule.f90:12-27: in-place filter of one plane, normalisation in G */
static void test_filter_plane(or_ctx *c, double *f, const double *G) {
  const lp_grid *g = &c->g;
  size_t nh = (size_t)g->lh * g->ny;
  cplx *F = (cplx *)malloc(sizeof(cplx) * nh);
  or_rfft2_fwd(c, 0, f, g->ld, (double *)F);
  for (size_t i = 0; i < nh; ++i) F[i] = G[i] * F[i];
  or_rfft2_inv(c, 0, (double *)F, f, g->ld);
  free(F);
}
void or_test_filter(or_ctx *c, double *f, const double *G, int nplanes) {
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nplanes; ++k) test_filter_plane(c, f + (size_t)k * plane(c), G);
}

/* ------------------------------------------------------------------------- */
/* convec.f90:9-316 (OCEAN/Stokes drift term zero: U_stokes = 0)             */
/* ------------------------------------------------------------------------- */
int or_convec(or_ctx *c) {
  const lp_grid *g = &c->g;
  int nz = g->nz;
  size_t np = plane(c), npb = (size_t)g->ld_big * g->ny2;
  for (int i = 0; i < 7; ++i)
    if (!c->big[i]) c->big[i] = (double *)calloc(npb * nz, sizeof(double));
  double *u1b = c->big[0], *u2b = c->big[1], *u3b = c->big[2];
  double *w1b = c->big[3], *w2b = c->big[4], *w3b = c->big[5], *ccb = c->big[6];
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  double *dudy = c->f[LP_DUDY], *dudz = c->f[LP_DUDZ], *dvdx = c->f[LP_DVDX];
  double *dvdz = c->f[LP_DVDZ], *dwdx = c->f[LP_DWDX], *dwdy = c->f[LP_DWDY];
  double *cx = c->f[LP_RHSX], *cy = c->f[LP_RHSY], *cz = c->f[LP_RHSZ];
  /* :76-98 velocities to the 3/2 grid */
  or_dealias1(c, u, u1b, nz);
  or_dealias1(c, v, u2b, nz);
  or_dealias1(c, w, u3b, nz);
  /* :100-133 vorticity (jz = 1 on uvp nodes) to the 3/2 grid */
  double *vx = (double *)calloc(np * nz, sizeof(double));
  double *vy = (double *)calloc(np * nz, sizeof(double));
  double *vz = (double *)calloc(np * nz, sizeof(double));
  for (int k = 0; k < nz; ++k)
    for (size_t i = k * np; i < (k + 1) * np; ++i) {
      if (k == 0) {
        vx[i] = 0.5 * (dwdy[i] + dwdy[i + np]) - dvdz[i];
        vy[i] = dudz[i] - 0.5 * (dwdx[i] + dwdx[i + np]);
      } else {
        vx[i] = dwdy[i] - dvdz[i];
        vy[i] = dudz[i] - dwdx[i];
      }
      vz[i] = dvdx[i] - dudy[i];
    }
  or_dealias1(c, vx, w1b, nz);
  or_dealias1(c, vy, w2b, nz);
  or_dealias1(c, vz, w3b, nz);
  free(vx); free(vy); free(vz);
  const double cb = 1.0 / ((double)g->nx2 * g->ny2);
  for (int comp = 0; comp < 3; ++comp) {
    double *out = comp == 0 ? cx : comp == 1 ? cy : cz;
    /* :140-158 (cx), :180-221 (cy), :241-279 (cz) products on the big grid */
#pragma omp parallel for
    for (int k = 0; k < nz - 1; ++k)
      for (size_t i = 0; i < npb; ++i) {
        size_t a = k * npb + i, b = a + npb, m = a - npb;
        double val;
        if (comp == 0) {
          if (k == 0) val = cb * (u2b[a] * (-w3b[a]) + 0.5 * u3b[b] * (w2b[b]));
          else val = cb * (u2b[a] * (-w3b[a]) + 0.5 * (u3b[b] * (w2b[b]) + u3b[a] * (w2b[a])));
        } else if (comp == 1) {
          if (k == 0) val = cb * (u1b[a] * (w3b[a]) + 0.5 * u3b[b] * (-w1b[b]));
          else val = cb * (u1b[a] * (w3b[a]) + 0.5 * (u3b[b] * (-w1b[b]) + u3b[a] * (-w1b[a])));
        } else {
          if (k == 0) val = 0.0;
          else val = cb * 0.5 * ((u1b[a] + u1b[m]) * (-w2b[a]) + (u2b[a] + u2b[m]) * (w1b[a]));
        }
        ccb[a] = val;
      }
    /* :161-170 etc: r2c big, unpadd, c2r small */
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nz - 1; ++k) {
      double *pb = ccb + k * npb;
      or_rfft2_fwd(c, 1, pb, g->ld_big, pb);
      cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)g->lh * g->ny);
      unpadd_plane(g, F, (const cplx *)pb);
      or_rfft2_inv(c, 0, (double *)F, out + k * np, g->ld);
      free(F);
    }
    for (size_t i = 0; i < np; ++i) out[(nz - 1) * np + i] = BOGUS; /* :306-308 */
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* wallstress.f90:5-215 (OCEAN_FLAG off, WALL_HOMOG_FLAG off; psi_m (plane)  */
/* and phi_m from obukhov, zero and one with theta off)                       */
/* ------------------------------------------------------------------------- */
int or_wallstress(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  size_t np = plane(c);
  double *u = c->f[LP_U], *v = c->f[LP_V];
  double *txz = c->f[LP_TXZ], *tyz = c->f[LP_TYZ], *dudz = c->f[LP_DUDZ], *dvdz = c->f[LP_DVDZ];
  double *ust = c->f[LP_USTAR];
  if (cf->lbc_mom == LP_LBC_STRESS_FREE) {                                  /* :200-205 */
    for (size_t i = 0; i < np; ++i) { txz[i] = 0; tyz[i] = 0; dudz[i] = 0; dvdz[i] = 0; }
    return 0;
  }
  if (cf->lbc_mom != LP_LBC_WALL) return fail(LP_ERR_ARG, "invalid lbc_mom");
  double *u1 = (double *)malloc(sizeof(double) * np), *v1 = (double *)malloc(sizeof(double) * np);
  memcpy(u1, u, sizeof(double) * np);                                       /* :103-104 */
  memcpy(v1, v, sizeof(double) * np);
  test_filter_plane(c, u1, c->G_test);                                      /* :139-140 */
  test_filter_plane(c, v1, c->G_test);
  /* psi_m and phi_m from obukhov: zero and one with theta off */
  const double *psi = c->f[LP_PSI_M];
  const double phi_m = cf->scal.s_flag ? c->obuk[2] : 1.0;
  double zl = 0.5 * g->dz - cf->d0;
  double denom = log(zl / cf->z0);                                          /* :143 */
  if (c->frz_ust) {
    /* frozen coefficient: txz = -(ustar^2 / |U|) u1 with ustar^2 / |U| = ustar kappa / denom
     * of the base run; dudz = ustar / (zl kappa) u / |U| = u / (zl denom) */
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = (size_t)y * g->ld + x;
        double q = -c->frz_ust[i] * cf->vonk / denom;
        txz[i] = q * u1[i];
        tyz[i] = q * v1[i];
        dudz[i] = u[i] / (zl * denom);
        dvdz[i] = v[i] / (zl * denom);
        ust[i] = c->frz_ust[i];
      }
    free(u1); free(v1);
    return 0;
  }
  for (int y = 0; y < g->ny; ++y)
    for (int x = 0; x < g->nx; ++x) {
      size_t i = (size_t)y * g->ld + x;
      double u_avg = sqrt(u1[i] * u1[i] + v1[i] * v1[i]);                   /* :145 */
      double us = u_avg * cf->vonk / (denom - psi[i]);                      /* :143, 162 */
      double cst = -(us * us) / u_avg;                                      /* :167 */
      txz[i] = cst * u1[i];
      tyz[i] = cst * v1[i];
      double du = us / (zl * cf->vonk) * u[i] / u_avg * phi_m;              /* :176-179 */
      double dv = us / (zl * cf->vonk) * v[i] / u_avg * phi_m;
      dudz[i] = (u[i] == 0.0) ? 0.0 : du;                                   /* :182-183 */
      dvdz[i] = (v[i] == 0.0) ? 0.0 : dv;
      ust[i] = us;
    }
  free(u1); free(v1);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* interpolag_Sdep.f90:2-373                                                  */
/* ------------------------------------------------------------------------- */
/* nf of the Lagrangian averages (fid) carried back along the trajectories.  ssim 0:
 * interpolag_Sdep.f90:2-373 (four averages; the wall's zp doubled and clamped to [-1, 1],
 * :281-291); ssim 1: interpolag_Ssim.f90:1-160 (F_LM, F_MM; at the wall only
 * zp(2) += min(zp(2), 0) and zp(1) += max(zp(1), 0), :176-185) */
static int interpolag(or_ctx *c, int nf, const int *fid, int ssim) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld;
  size_t np = plane(c);
  size_t fn = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  double *FF[4];
#define FFI(i, j, k) (((size_t)(k) * (ny + 2) + (j)) * (nx + 2) + (i))
  for (int q = 0; q < nf; ++q) {
    FF[q] = (double *)malloc(sizeof(double) * fn);
    const double *F = c->f[fid[q]];
    /* :62-71 interior; :75-105 periodic x, y wrap and z copy (incl. corners) */
    for (int k = 0; k < nz + 2; ++k) {
      int ks = k - 1; if (ks < 0) ks = 0; if (ks > nz - 1) ks = nz - 1;
      for (int j = 0; j < ny + 2; ++j) {
        int js = (j - 1 + ny) % ny;
        for (int i = 0; i < nx + 2; ++i) {
          int is = (i - 1 + nx) % nx;
          FF[q][FFI(i, j, k)] = F[(size_t)ks * np + (size_t)js * ld + is];
        }
      }
    }
  }
  double cs = (double)cf->cs_count;
  double *ul = c->f[LP_U_LAG], *vl = c->f[LP_V_LAG], *wl = c->f[LP_W_LAG];
  double *ut = (double *)malloc(sizeof(double) * np * nz), *vt = (double *)malloc(sizeof(double) * np * nz);
  for (size_t i = 0; i < np * nz; ++i) { ut[i] = ul[i] / cs; vt[i] = vl[i] / cs; }  /* :204-205 */
  for (size_t i = 0; i < np; ++i) { ul[i] = ut[i]; vl[i] = vt[i]; }                   /* :215,226 */
  for (int k = 1; k < nz; ++k)                                                         /* :235-248 */
    for (size_t i = 0; i < np; ++i) {
      ul[k * np + i] = 0.5 * (ut[k * np + i] + ut[(k - 1) * np + i]);
      vl[k * np + i] = 0.5 * (vt[k * np + i] + vt[(k - 1) * np + i]);
    }
  for (size_t i = 0; i < np * nz; ++i) wl[i] = wl[i] / cs;                             /* :250 */
  for (size_t i = 0; i < np; ++i) wl[i] = 0.25 * wl[np + i];                          /* :258 */
  free(ut); free(vt);
  double *xp = (double *)malloc(sizeof(double) * np * nz);
  double *yp = (double *)malloc(sizeof(double) * np * nz);
  double *zp = (double *)malloc(sizeof(double) * np * nz);
  for (size_t i = 0; i < np * nz; ++i) {                                               /* :263-265 */
    xp[i] = -ul[i] * cf->dt * cs / g->dx;
    yp[i] = -vl[i] * cf->dt * cs / g->dy;
    zp[i] = -wl[i] * cf->dt * cs / g->dz;
  }
  for (size_t i = 0; i < np; ++i) {
    double z2 = zp[np + i];
    z2 = z2 + fmin(z2, 0.0);
    if (ssim) {                                                                        /* Ssim :180-185 */
      zp[np + i] = z2;
      zp[i] = zp[i] + fmax(zp[i], 0.0);
    } else {                                                                           /* Sdep :281-291 */
      z2 = fmin(1.0, z2);
      z2 = fmax(-1.0, z2);
      zp[np + i] = z2;
      zp[i] = z2;
    }
  }
  double *Fo[4];
  for (int q = 0; q < nf; ++q) Fo[q] = c->f[fid[q]];
#pragma omp parallel for
  for (int k = 0; k < nz; ++k)                                                          /* :301-365 */
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t s = (size_t)k * np + (size_t)j * ld + i;
        int jjx = i + 1, jjy = j + 1, jjz = k + 1;
        int addx = signbit(xp[s]) ? -1 : 1;
        int addy = signbit(yp[s]) ? -1 : 1;
        int addz = signbit(zp[s]) ? -1 : 1;
        int jxa = jjx + addx, jya = jjy + addy, jza = jjz + addz;
        double cx_ = fabs(xp[s]), cy_ = fabs(yp[s]), cz_ = fabs(zp[s]);
        double fx = 1.0 - cx_, fy = 1.0 - cy_, fz = 1.0 - cz_;
        for (int q = 0; q < nf; ++q) {
          const double *F = FF[q];
          Fo[q][s] = fx * fy * (F[FFI(jjx, jjy, jjz)] * fz + F[FFI(jjx, jjy, jza)] * cz_)
                   + fx * cy_ * (F[FFI(jjx, jya, jjz)] * fz + F[FFI(jjx, jya, jza)] * cz_)
                   + cx_ * fy * (F[FFI(jxa, jjy, jjz)] * fz + F[FFI(jxa, jjy, jza)] * cz_)
                   + cx_ * cy_ * (F[FFI(jxa, jya, jjz)] * fz + F[FFI(jxa, jya, jza)] * cz_);
        }
      }
#undef FFI
  for (size_t i = 0; i < np * nz; ++i) { ul[i] = 0; vl[i] = 0; wl[i] = 0; }           /* :367-369 */
  for (int q = 0; q < nf; ++q) free(FF[q]);
  free(xp); free(yp); free(zp);
  return 0;
}
int or_interpolag_sdep(or_ctx *c) {
  const int fid[4] = {LP_F_LM, LP_F_MM, LP_F_QN, LP_F_NN};
  return interpolag(c, 4, fid, 0);
}
int or_interpolag_ssim(or_ctx *c) {
  const int fid[2] = {LP_F_LM, LP_F_MM};
  return interpolag(c, 2, fid, 1);
}

/* ------------------------------------------------------------------------- */
/* lagrange_Sdep.f90:4-496 (use_bldg, inflow off)                             */
/* ------------------------------------------------------------------------- */
static inline double r4(double x) { return (double)(float)x; }

int or_lagrange_sdep(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld;
  size_t np = plane(c);
  const int casts = cf->fp32_casts;
  const double tiny = casts ? (double)1e-24f : 1e-24;
  double delta = cf->filter_size * pow(g->dx * g->dy * g->dz, 1.0 / 3.0);  /* :75 */
  const double tf1 = 2.0, tf2 = 4.0, tf1_2 = tf1 * tf1, tf2_2 = tf2 * tf2;
  double opftdelta = cf->opftime * delta;
  const double powcoeff = -1.0 / 8.0;
  int rc = or_interpolag_sdep(c);                                            /* :120 */
  if (rc) return rc;
  double lagran_dt = cf->dt * (double)cf->cs_count;                          /* :123 */
  const double cst = 2.0 * delta * delta;                                    /* :125 */
  const int do_init_lm = cf->inilag && !c->F_LM_MM_init &&
                         (c->jt == cf->cs_count || c->jt == cf->dyn_init);   /* :277-278 */
  const int do_init_qn = cf->inilag && !c->F_QN_NN_init &&
                         (c->jt == cf->cs_count || c->jt == cf->dyn_init);   /* :365-366 */
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  double *F_LM = c->f[LP_F_LM], *F_MM = c->f[LP_F_MM], *F_QN = c->f[LP_F_QN], *F_NN = c->f[LP_F_NN];
  double *Cs = c->f[LP_CS_OPT2], *Beta = c->f[LP_BETA];
  double **S = c->S;
  const double bexp = log(tf1) / (log(tf2) - log(tf1));
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nz; ++k) {
    double *buf = (double *)malloc(sizeof(double) * np * 45);
    double *ub = buf, *vb = buf + np, *wb = buf + 2 * np;
    double *uh = buf + 3 * np, *vh = buf + 4 * np, *wh = buf + 5 * np;
    double *L[6], *Q[6], *Sb[6], *Sh[6], *SSb[6], *SSh[6];
    for (int q = 0; q < 6; ++q) {
      L[q] = buf + (6 + q) * np; Q[q] = buf + (12 + q) * np;
      Sb[q] = buf + (18 + q) * np; Sh[q] = buf + (24 + q) * np;
      SSb[q] = buf + (30 + q) * np; SSh[q] = buf + (36 + q) * np;
    }
    double *Sm = buf + 42 * np;
    size_t o = (size_t)k * np;
    for (size_t i = 0; i < np; ++i) {                                          /* :129-140 */
      if (k == 0) { ub[i] = u[i]; vb[i] = v[i]; wb[i] = 0.25 * w[np + i]; }
      else {
        ub[i] = 0.5 * (u[o + i] + u[o - np + i]);
        vb[i] = 0.5 * (v[o + i] + v[o - np + i]);
        wb[i] = w[o + i];
      }
      uh[i] = ub[i]; vh[i] = vb[i]; wh[i] = wb[i];                               /* :141-143 */
      L[0][i] = ub[i] * ub[i]; L[1][i] = ub[i] * vb[i]; L[2][i] = ub[i] * wb[i];  /* :144-149 */
      L[3][i] = vb[i] * vb[i]; L[4][i] = vb[i] * wb[i]; L[5][i] = wb[i] * wb[i];
      for (int q = 0; q < 6; ++q) Q[q][i] = L[q][i];                              /* :150-155 */
    }
    /* component order q: 0=11 1=12 2=13 3=22 4=23 5=33 */
    test_filter_plane(c, ub, c->G_test); test_filter_plane(c, vb, c->G_test);
    test_filter_plane(c, wb, c->G_test);
    for (int q = 0; q < 6; ++q) test_filter_plane(c, L[q], c->G_test);
    test_filter_plane(c, uh, c->G_test_test); test_filter_plane(c, vh, c->G_test_test);
    test_filter_plane(c, wh, c->G_test_test);
    for (int q = 0; q < 6; ++q) test_filter_plane(c, Q[q], c->G_test_test);
    for (size_t i = 0; i < np; ++i) {                                          /* :162-189 */
      L[0][i] -= ub[i] * ub[i]; L[1][i] -= ub[i] * vb[i]; L[2][i] -= ub[i] * wb[i];
      L[3][i] -= vb[i] * vb[i]; L[4][i] -= vb[i] * wb[i]; L[5][i] -= wb[i] * wb[i];
      Q[0][i] -= uh[i] * uh[i]; Q[1][i] -= uh[i] * vh[i]; Q[2][i] -= uh[i] * wh[i];
      Q[3][i] -= vh[i] * vh[i]; Q[4][i] -= vh[i] * wh[i]; Q[5][i] -= wh[i] * wh[i];
    }
    for (size_t i = 0; i < np; ++i) {                                          /* :192-242 */
      double s11 = S[0][o + i], s12 = S[1][o + i], s13 = S[2][o + i];
      double s22 = S[3][o + i], s23 = S[4][o + i], s33 = S[5][o + i];
      double sm = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 +
                              2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
      Sm[i] = sm;
      for (int q = 0; q < 6; ++q) {
        Sb[q][i] = S[q][o + i];
        Sh[q][i] = S[q][o + i];
        SSb[q][i] = sm * S[q][o + i];
        SSh[q][i] = SSb[q][i];
      }
    }
    for (int q = 0; q < 6; ++q) test_filter_plane(c, Sb[q], c->G_test);          /* :210-222 */
    for (int q = 0; q < 6; ++q) test_filter_plane(c, Sh[q], c->G_test_test);
    for (int q = 0; q < 6; ++q) test_filter_plane(c, SSb[q], c->G_test);        /* :244-256 */
    for (int q = 0; q < 6; ++q) test_filter_plane(c, SSh[q], c->G_test_test);
    double *FLM = F_LM + o, *FMM = F_MM + o, *FQN = F_QN + o, *FNN = F_NN + o;
    double *cso = Cs + o, *bet = Beta + o;
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < ld; ++x) {
        size_t i = (size_t)y * ld + x;
        double sbar = sqrt(2.0 * (Sb[0][i] * Sb[0][i] + Sb[3][i] * Sb[3][i] + Sb[5][i] * Sb[5][i] +
                                  2.0 * (Sb[1][i] * Sb[1][i] + Sb[2][i] * Sb[2][i] + Sb[4][i] * Sb[4][i])));
        double shat = sqrt(2.0 * (Sh[0][i] * Sh[0][i] + Sh[3][i] * Sh[3][i] + Sh[5][i] * Sh[5][i] +
                                  2.0 * (Sh[1][i] * Sh[1][i] + Sh[2][i] * Sh[2][i] + Sh[4][i] * Sh[4][i])));
        double M[6], N[6];
        for (int q = 0; q < 6; ++q) {                                          /* :258-270 */
          M[q] = cst * (SSb[q][i] - tf1_2 * sbar * Sb[q][i]);
          N[q] = cst * (SSh[q][i] - tf2_2 * shat * Sh[q][i]);
        }
        double LM = L[0][i] * M[0] + L[3][i] * M[3] + L[5][i] * M[5] +           /* :272-275 */
                    2.0 * (L[1][i] * M[1] + L[2][i] * M[2] + L[4][i] * M[4]);
        double MM = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] +
                    2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
        double QN = Q[0][i] * N[0] + Q[3][i] * N[3] + Q[5][i] * N[5] +
                    2.0 * (Q[1][i] * N[1] + Q[2][i] * N[2] + Q[4][i] * N[4]);
        double NN = N[0] * N[0] + N[3] * N[3] + N[5] * N[5] +
                    2.0 * (N[1] * N[1] + N[2] * N[2] + N[4] * N[4]);
        int padc = (x >= ld - 2);
        if (do_init_lm) {                                                       /* :279-290 */
          FMM[i] = MM;
          FLM[i] = 0.0;
          if (padc) { FMM[i] = 1.0; FLM[i] = 1.0; }
        }
        double Tn, dum, epsi;
        if (casts) {                                                            /* :311-315 */
          Tn = (double)fmaxf((float)(FLM[i] * FMM[i]), 1e-24f);
          Tn = opftdelta * pow(Tn, powcoeff);
          Tn = (double)fmaxf(1e-24f, (float)Tn);
        } else {
          Tn = fmax(FLM[i] * FMM[i], tiny);
          Tn = opftdelta * pow(Tn, powcoeff);
          Tn = fmax(tiny, Tn);
        }
        dum = lagran_dt / Tn;
        epsi = dum / (1.0 + dum);
        FLM[i] = (epsi * LM + (1.0 - epsi) * FLM[i]);                          /* :324-325 */
        FMM[i] = (epsi * MM + (1.0 - epsi) * FMM[i]);
        FLM[i] = casts ? (double)fmaxf(1e-24f, (float)FLM[i]) : fmax(tiny, FLM[i]); /* :329 */
        double cs2d = FLM[i] / FMM[i];                                           /* :358-363 */
        if (padc) cs2d = tiny;
        cs2d = casts ? r4(fmax(tiny, r4(cs2d))) : fmax(tiny, cs2d);
        if (do_init_qn) {                                                       /* :366-379 */
          FNN[i] = NN;
          FQN[i] = 0.03 * NN;
          if (padc) { FNN[i] = 1.0; FQN[i] = 1.0; }
        }
        if (casts) {                                                            /* :386-390 */
          Tn = (double)fmaxf((float)(FQN[i] * FNN[i]), 1e-24f);
          Tn = opftdelta * pow(Tn, powcoeff);
          Tn = (double)fmaxf(1e-24f, (float)Tn);
        } else {
          Tn = fmax(FQN[i] * FNN[i], tiny);
          Tn = opftdelta * pow(Tn, powcoeff);
          Tn = fmax(tiny, Tn);
        }
        dum = lagran_dt / Tn;
        epsi = dum / (1.0 + dum);
        FQN[i] = (epsi * QN + (1.0 - epsi) * FQN[i]);                          /* :392-393 */
        FNN[i] = (epsi * NN + (1.0 - epsi) * FNN[i]);
        FQN[i] = casts ? (double)fmaxf(1e-24f, (float)FQN[i]) : fmax(tiny, FQN[i]); /* :397 */
        double cs4d = FQN[i] / FNN[i];                                           /* :403-407 */
        if (padc) cs4d = tiny;
        cs4d = casts ? r4(fmax(tiny, r4(cs4d))) : fmax(tiny, cs4d);
        double b = pow(cs4d / cs2d, bexp);                                       /* :409-410 */
        if (k == nz - 1) b = 1.0;                                                /* :425-428 */
        bet[i] = b;
        double bclip = casts ? (double)fmaxf((float)b, (float)(1.0 / (tf1 * tf2)))  /* :430 */
                             : fmax(b, 1.0 / (tf1 * tf2));
        double csv = cs2d / bclip;                                               /* :431-434 */
        if (padc) csv = tiny;
        csv = casts ? (double)fmaxf(1e-24f, (float)csv) : fmax(tiny, csv);
        cso[i] = csv;
      }
    free(buf);
  }
  if (do_init_lm) c->F_LM_MM_init = 1;                                        /* :292 */
  if (do_init_qn) c->F_QN_NN_init = 1;                                        /* :380 */
  return 0;
}

/* ------------------------------------------------------------------------- */
/* The other dynamic SGS models of sgs_stag.f90:188-198: std_dynamic.f90      */
/* (model 2), scaledep_dynamic.f90 (3), lagrange_Ssim.f90 (4).  Their plane   */
/* sums run over the physical columns x < nx (the reference sums the (ld, ny) */
/* planes, whose padding columns hold zero after the in-place inverse of a    */
/* spectrum with no Nyquist column).                                          */
/* ------------------------------------------------------------------------- */
/* The filtered fields of one level: 2-delta (G_test) L_ij, S_ij, |S| S_ij, |S|bar;
 * with hat, also the 4-delta ones of scaledep_dynamic.f90:85-152 (Q_ij from the
 * products of the 2-delta filtered velocities, :100-120). Component order 11 12 13 22 23 33. */
typedef struct {
  double *L[6], *Sb[6], *SSb[6], *sbar;
  double *Q[6], *Sh[6], *SSh[6], *shat;
  double *buf;
} dyn_lev;

static void dyn_level(or_ctx *c, int k, int hat, dyn_lev *d) {
  const size_t np = plane(c);
  const size_t o = (size_t)k * np;
  const double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  double **S = c->S;
  d->buf = (double *)calloc(np * 48, sizeof(double));
  double *b = d->buf;
  double *ub = b, *vb = b + np, *wb = b + 2 * np, *uh = b + 3 * np, *vh = b + 4 * np, *wh = b + 5 * np;
  for (int q = 0; q < 6; ++q) {
    d->L[q] = b + (6 + q) * np; d->Q[q] = b + (12 + q) * np;
    d->Sb[q] = b + (18 + q) * np; d->Sh[q] = b + (24 + q) * np;
    d->SSb[q] = b + (30 + q) * np; d->SSh[q] = b + (36 + q) * np;
  }
  d->sbar = b + 42 * np; d->shat = b + 43 * np;
  double *Sm = b + 44 * np;
  for (size_t i = 0; i < np; ++i) {                                /* std_dynamic.f90:36-62 */
    if (k == 0) { ub[i] = u[i]; vb[i] = v[i]; wb[i] = 0.25 * w[np + i]; }
    else {
      ub[i] = 0.5 * (u[o + i] + u[o - np + i]);
      vb[i] = 0.5 * (v[o + i] + v[o - np + i]);
      wb[i] = w[o + i];
    }
    uh[i] = ub[i]; vh[i] = vb[i]; wh[i] = wb[i];
    d->L[0][i] = ub[i] * ub[i]; d->L[1][i] = ub[i] * vb[i]; d->L[2][i] = ub[i] * wb[i];
    d->L[3][i] = vb[i] * vb[i]; d->L[4][i] = vb[i] * wb[i]; d->L[5][i] = wb[i] * wb[i];
  }
  test_filter_plane(c, ub, c->G_test); test_filter_plane(c, vb, c->G_test);   /* :64-78 */
  test_filter_plane(c, wb, c->G_test);
  for (int q = 0; q < 6; ++q) test_filter_plane(c, d->L[q], c->G_test);
  for (size_t i = 0; i < np; ++i) {
    d->L[0][i] -= ub[i] * ub[i]; d->L[1][i] -= ub[i] * vb[i]; d->L[2][i] -= ub[i] * wb[i];
    d->L[3][i] -= vb[i] * vb[i]; d->L[4][i] -= vb[i] * wb[i]; d->L[5][i] -= wb[i] * wb[i];
  }
  if (hat) {                                                       /* scaledep :100-120 */
    for (size_t i = 0; i < np; ++i) {
      d->Q[0][i] = ub[i] * ub[i]; d->Q[1][i] = ub[i] * vb[i]; d->Q[2][i] = ub[i] * wb[i];
      d->Q[3][i] = vb[i] * vb[i]; d->Q[4][i] = vb[i] * wb[i]; d->Q[5][i] = wb[i] * wb[i];
    }
    test_filter_plane(c, uh, c->G_test_test); test_filter_plane(c, vh, c->G_test_test);
    test_filter_plane(c, wh, c->G_test_test);
    for (int q = 0; q < 6; ++q) test_filter_plane(c, d->Q[q], c->G_test_test);
    for (size_t i = 0; i < np; ++i) {
      d->Q[0][i] -= uh[i] * uh[i]; d->Q[1][i] -= uh[i] * vh[i]; d->Q[2][i] -= uh[i] * wh[i];
      d->Q[3][i] -= vh[i] * vh[i]; d->Q[4][i] -= vh[i] * wh[i]; d->Q[5][i] -= wh[i] * wh[i];
    }
  }
  for (size_t i = 0; i < np; ++i) {                                /* std_dynamic.f90:80-110 */
    double s11 = S[0][o + i], s12 = S[1][o + i], s13 = S[2][o + i];
    double s22 = S[3][o + i], s23 = S[4][o + i], s33 = S[5][o + i];
    Sm[i] = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
    for (int q = 0; q < 6; ++q) {
      d->Sb[q][i] = S[q][o + i];
      d->Sh[q][i] = S[q][o + i];
      d->SSb[q][i] = Sm[i] * S[q][o + i];
      d->SSh[q][i] = d->SSb[q][i];
    }
  }
  for (int q = 0; q < 6; ++q) {
    test_filter_plane(c, d->Sb[q], c->G_test);
    test_filter_plane(c, d->SSb[q], c->G_test);
    if (hat) {
      test_filter_plane(c, d->Sh[q], c->G_test_test);
      test_filter_plane(c, d->SSh[q], c->G_test_test);
    }
  }
#define MAG(A, i) sqrt(2.0 * (A[0][i] * A[0][i] + A[3][i] * A[3][i] + A[5][i] * A[5][i] + \
                              2.0 * (A[1][i] * A[1][i] + A[2][i] * A[2][i] + A[4][i] * A[4][i])))
  for (size_t i = 0; i < np; ++i) {
    d->sbar[i] = MAG(d->Sb, i);
    if (hat) d->shat[i] = MAG(d->Sh, i);
  }
#undef MAG
}

/* a . b in the reference's order: 11 + 22 + 33 + 2 (12 + 13 + 23) */
static inline double dot6(double *const *a, double *const *b, size_t i) {
  return a[0][i] * b[0][i] + a[3][i] * b[3][i] + a[5][i] * b[5][i] +
         2.0 * (a[1][i] * b[1][i] + a[2][i] * b[2][i] + a[4][i] * b[4][i]);
}

/* L.M and M.M at one point with M_ij = cst (|S|S_ij bar - 4 b |S|bar S_ij bar) */
static void lm_mm(const dyn_lev *d, size_t i, double cst, double fourb, double *lm, double *mm) {
  double M[6];
  for (int q = 0; q < 6; ++q) M[q] = cst * (d->SSb[q][i] - fourb * d->sbar[i] * d->Sb[q][i]);
  *lm = d->L[0][i] * M[0] + d->L[3][i] * M[3] + d->L[5][i] * M[5] +
        2.0 * (d->L[1][i] * M[1] + d->L[2][i] * M[2] + d->L[4][i] * M[4]);
  *mm = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
}

/* sgsmodule.f90:47-69 rtnewt: Newton from 7.5 on the quintic A, |dx| < 0.001, else 1 */
static double rtnewt(const double *A) {
  double r = 0.5 * (0.0 + 15.0);
  for (int j = 0; j < 100; ++j) {
    const double f = A[0] + r * (A[1] + r * (A[2] + r * (A[3] + r * (A[4] + r * A[5]))));
    const double df = A[1] + r * (2.0 * A[2] + r * (3.0 * A[3] + r * (4.0 * A[4] + r * (5.0 * A[5]))));
    const double dx = f / df;
    r = r - dx;
    if (fabs(dx) < 0.001) return r;
  }
  return 1.0;
}

/* std_dynamic.f90:1-136 (model 2) and scaledep_dynamic.f90:1-230 (model 3): one Cs^2 per
 * level from plane sums, written over the whole plane (sgs_stag.f90:190-194).  Model 3's
 * beta, local to scaledep_dynamic (its output file, :231-245, is not written), is left in
 * Beta, which no other part of these models reads */
/* one level: out[0..9] = a1, b1, c1, d1, e1, a2, b2, c2, d2, e2 (scaledep_dynamic.f90:156-198;
 * the 4-delta ones only for model 3), out[10] = beta, out[11] = Cs^2 */
static void plane_dynamic_level(or_ctx *c, int model, int k, double *out) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  const int nx = g->nx, ny = g->ny, ld = g->ld;
  const double delta = cf->filter_size * pow(g->dx * g->dy * g->dz, 1.0 / 3.0);
  const double cst = 2.0 * delta * delta;
  const double nxy = (double)(nx * ny);
  {
    dyn_lev d;
    dyn_level(c, k, model == 3, &d);
    double beta = 1.0;
    {                                                              /* scaledep :156-209 */
      double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const size_t i = (size_t)y * ld + x;
          const double sb = d.sbar[i], sh = d.shat[i];
          s[0] += sb * dot6(d.Sb, d.L, i);
          s[1] += dot6(d.SSb, d.L, i);
          s[2] += dot6(d.SSb, d.SSb, i);
          s[3] += 0.5 * (sb * sb) * (sb * sb);
          s[4] += sb * dot6(d.Sb, d.SSb, i);
          s[5] += sh * dot6(d.Sh, d.Q, i);
          s[6] += dot6(d.SSh, d.Q, i);
          s[7] += dot6(d.SSh, d.SSh, i);
          s[8] += 0.5 * (sh * sh) * (sh * sh);
          s[9] += sh * dot6(d.Sh, d.SSh, i);
        }
      const double d2 = delta * delta, c2 = (2.0 * d2) * (2.0 * d2);
      const double a1 = -2.0 * d2 * 4.0 * s[0] / nxy, b1 = -2.0 * d2 * s[1] / nxy;
      const double c1 = c2 * s[2] / nxy, dd1 = c2 * 16.0 * s[3] / nxy, e1 = 2.0 * c2 * 4.0 * s[4] / nxy;
      const double a2 = -2.0 * d2 * 16.0 * s[5] / nxy, b2 = -2.0 * d2 * s[6] / nxy;
      const double cc2 = c2 * s[7] / nxy, dd2 = c2 * 256.0 * s[8] / nxy, e2 = 2.0 * c2 * 16.0 * s[9] / nxy;
      double A[6];
      A[0] = b2 * c1 - b1 * cc2;
      A[1] = a1 * cc2 - b2 * e1;
      A[2] = b2 * dd1 + b1 * e2 - a2 * c1;
      A[3] = a2 * e1 - a1 * e2;
      A[4] = -a2 * dd1 - b1 * dd2;
      A[5] = a1 * dd2;
      if (model == 3) beta = rtnewt(A);
      const double ab[10] = {a1, b1, c1, dd1, e1, a2, b2, cc2, dd2, e2};
      memcpy(out, ab, sizeof ab);
    }
    double slm = 0.0, smm = 0.0;                                   /* std_dynamic :126-130 */
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        double lm, mm;
        lm_mm(&d, (size_t)y * ld + x, cst, 4.0 * beta, &lm, &mm);
        slm += lm;
        smm += mm;
      }
    double cs = slm / smm;
    cs = fmax(0.0, cs);
    out[10] = beta;
    out[11] = cs;
    free(d.buf);
  }
}
static int or_plane_dynamic(or_ctx *c, int model) {
  const size_t np = plane(c);
  const int nz = c->g.nz;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nz; ++k) {
    double out[12];
    plane_dynamic_level(c, model, k, out);
    for (size_t i = 0; i < np; ++i) {
      c->f[LP_CS_OPT2][(size_t)k * np + i] = out[11];
      if (model == 3) c->f[LP_BETA][(size_t)k * np + i] = out[10];
    }
  }
  return 0;
}
/* L.M and M.M (M with 4 |S|bar S_ij bar) at every point of level k (plane arrays; zero in
 * the padding columns) from the current velocities and strain (after an or_sgs_stag) */
void or_dyn_contract(or_ctx *c, int k, double *lm, double *mm) {
  const lp_grid *g = &c->g;
  const double delta = c->cfg.filter_size * pow(g->dx * g->dy * g->dz, 1.0 / 3.0);
  dyn_lev d;
  dyn_level(c, k, 0, &d);
  for (int y = 0; y < g->ny; ++y)
    for (int x = 0; x < g->ld; ++x) {
      const size_t i = (size_t)y * g->ld + x;
      lm[i] = mm[i] = 0.0;
      if (x < g->nx) lm_mm(&d, i, 2.0 * delta * delta, 4.0, &lm[i], &mm[i]);
    }
  free(d.buf);
}
/* the sums, beta and Cs^2 of level k from the current strain (after an or_sgs_stag) */
void or_plane_dynamic_level(or_ctx *c, int model, int k, double *out) { plane_dynamic_level(c, model, k, out); }
int or_std_dynamic(or_ctx *c) { return or_plane_dynamic(c, 2); }
int or_scaledep_dynamic(or_ctx *c) { return or_plane_dynamic(c, 3); }

/* lagrange_Ssim.f90:1-322 (model 4; use_bldg, TREES, inflow off: Beta = 1), with
 * interpolag_Ssim; the reference's REAL(4) conversions of :281-303 when fp32_casts */
int or_lagrange_ssim(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  const int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld;
  const size_t np = plane(c);
  const int casts = cf->fp32_casts;
  const double tiny = casts ? (double)1e-32f : 1e-32;
  const double delta = cf->filter_size * pow(g->dx * g->dy * g->dz, 1.0 / 3.0);
  const double opftdelta = cf->opftime * delta, powcoeff = -1.0 / 8.0;
  int rc = or_interpolag_ssim(c);                                          /* :103 */
  if (rc) return rc;
  const double lagran_dt = cf->dt * (double)cf->cs_count;
  const double cst = 2.0 * delta * delta;
  const int do_init = cf->inilag && !c->F_LM_MM_init && (c->jt == cf->cs_count || c->jt == cf->dyn_init);
  double *FLM = c->f[LP_F_LM], *FMM = c->f[LP_F_MM], *Cs = c->f[LP_CS_OPT2], *Bt = c->f[LP_BETA];
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < nz; ++k) {
    dyn_lev d;
    dyn_level(c, k, 0, &d);
    const size_t o = (size_t)k * np;
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < ld; ++x) {
        const size_t i = (size_t)y * ld + x, s = o + i;
        const int padc = x >= nx;
        double lm = 0.0, mm = 0.0;  /* padding columns: the zero Nyquist column */
        if (!padc) lm_mm(&d, i, cst, 4.0 * 1.0, &lm, &mm);               /* :178-192 */
        if (do_init) {                                                     /* :193-201 */
          FMM[s] = mm;
          FLM[s] = 0.025 * mm;
          if (x >= ld - 2) { FMM[s] = 1.0; FLM[s] = 1.0; }
        }
        double Tn = casts ? (double)fmaxf((float)(FLM[s] * FMM[s]), 1e-32f) : fmax(FLM[s] * FMM[s], tiny); /* :208 */
        Tn = opftdelta * pow(Tn, powcoeff);
        const double dum = lagran_dt / Tn, epsi = dum / (1.0 + dum);
        FLM[s] = (epsi * lm + (1.0 - epsi) * FLM[s]);                     /* :212-213 */
        FMM[s] = (epsi * mm + (1.0 - epsi) * FMM[s]);
        FLM[s] = casts ? (double)fmaxf(1e-32f, (float)FLM[s]) : fmax(tiny, FLM[s]);  /* :214 */
        double cs = FLM[s] / FMM[s];                                       /* :215-218 */
        if (x >= ld - 2) cs = 0.0;
        Cs[s] = casts ? (double)fmaxf(1e-32f, (float)cs) : fmax(tiny, cs);
        Bt[s] = 1.0;
      }
    free(d.buf);
  }
  if (do_init) c->F_LM_MM_init = 1;
  return 0;
}


/* ------------------------------------------------------------------------- */
/* sgs_stag.f90:3-401 (sgs on, molec off, OCEAN/SEABED off)                   */
/* ------------------------------------------------------------------------- */
int or_sgs_stag(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld;
  size_t np = plane(c), nv = vol(c);
  for (int q = 0; q < 6; ++q)
    if (!c->S[q]) c->S[q] = (double *)calloc(nv, sizeof(double));
  double **S = c->S;
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  double *dudx = c->f[LP_DUDX], *dudy = c->f[LP_DUDY], *dudz = c->f[LP_DUDZ];
  double *dvdx = c->f[LP_DVDX], *dvdy = c->f[LP_DVDY], *dvdz = c->f[LP_DVDZ];
  double *dwdx = c->f[LP_DWDX], *dwdy = c->f[LP_DWDY], *dwdz = c->f[LP_DWDZ];
  double *txx = c->f[LP_TXX], *txy = c->f[LP_TXY], *txz = c->f[LP_TXZ];
  double *tyy = c->f[LP_TYY], *tyz = c->f[LP_TYZ], *tzz = c->f[LP_TZZ];
  double *Cs = c->f[LP_CS_OPT2], *Nu_t = c->f[LP_NU_T], *dissip = c->f[LP_DISSIP], *magS = c->f[LP_MAGS];
  double delta = cf->filter_size * pow(g->dx * g->dy * g->dz, 1.0 / 3.0);    /* :46 */
  /* :49-78 wall level on uvp nodes */
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      size_t i = (size_t)y * ld + x;
      double ux = dudx[i], uy = dudy[i], uz = dudz[i];
      double vx = dvdx[i], vy = dvdy[i], vz = dvdz[i];
      double wx = 0.5 * (dwdx[i] + dwdx[i + np]);
      double wy = 0.5 * (dwdy[i] + dwdy[i + np]);
      double wz = dwdz[i];
      S[0][i] = ux; S[1][i] = 0.5 * (uy + vx); S[2][i] = 0.5 * (uz + wx);
      S[3][i] = vy; S[4][i] = 0.5 * (vz + wy); S[5][i] = wz;
    }
  /* :105-125 w-nodes */
#pragma omp parallel for
  for (int k = 1; k < nz; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x, m = i - np;
        double ux = 0.5 * (dudx[i] + dudx[m]), uy = 0.5 * (dudy[i] + dudy[m]), uz = dudz[i];
        double vx = 0.5 * (dvdx[i] + dvdx[m]), vy = 0.5 * (dvdy[i] + dvdy[m]), vz = dvdz[i];
        double wx = dwdx[i], wy = dwdy[i], wz = 0.5 * (dwdz[i] + dwdz[m]);
        S[0][i] = ux; S[1][i] = 0.5 * (uy + vx); S[2][i] = 0.5 * (uz + wx);
        S[3][i] = vy; S[4][i] = 0.5 * (vz + wy); S[5][i] = wz;
      }
  double *lz = (double *)malloc(sizeof(double) * nz);
  if (c->frz_nut) {
    /* frozen coefficient: Nu_t of the base run (no model update, no lag accumulation) */
    for (size_t i = 0; i < nv; ++i) Nu_t[i] = c->frz_nut[i];
    goto stresses;
  }
  if (cf->model == 4 || cf->model == 5) {                                     /* :130-134 */
    double *ul = c->f[LP_U_LAG], *vl = c->f[LP_V_LAG], *wl = c->f[LP_W_LAG];
    for (size_t i = 0; i < nv; ++i) { ul[i] += u[i]; vl[i] += v[i]; wl[i] += w[i]; }
  }
  if (cf->model == 1) {                                                       /* :137-174 */
    for (size_t i = 0; i < nv; ++i) Cs[i] = cf->Co * cf->Co;
    for (int k = 0; k < nz; ++k) {
      double zz = (k == 0) ? 0.5 * g->dz : (double)k * g->dz;
      double n = (double)cf->nnn;
      lz[k] = pow(pow(cf->Co, n) * pow(cf->vonk * zz, -n) + pow(delta, -n), -1.0 / n);
    }
  } else if (cf->model >= 2 && cf->model <= 5) {
    for (int k = 0; k < nz; ++k) lz[k] = delta;                               /* :177 */
    if (c->jt == 1 && cf->inilag) {                                           /* :178-180 */
      for (size_t i = 0; i < nv; ++i) Cs[i] = cf->cs2_init;
    } else if ((c->jt >= cf->dyn_init || cf->initu) && (c->jt % cf->cs_count) == 0) {
      int rc = cf->model == 5   ? or_lagrange_sdep(c)                         /* :200-201 */
               : cf->model == 4 ? or_lagrange_ssim(c)                         /* :198-199 */
                                : or_plane_dynamic(c, cf->model);             /* :188-197 */
      if (rc) { free(lz); return rc; }
    }
  } else {
    free(lz);
    return fail(LP_ERR_ARG, "unsupported SGS model (1 to 5 are implemented)");
  }
  /* :215-237 |S|, Nu_t, dissip, magS on 1..nz */
#pragma omp parallel for
  for (int k = 0; k < nz; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x;
        double s11 = S[0][i], s12 = S[1][i], s13 = S[2][i], s22 = S[3][i], s23 = S[4][i], s33 = S[5][i];
        double s = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
        double l2 = lz[k] * lz[k];
        Nu_t[i] = s * Cs[i] * l2;
        dissip[i] = -(s * s * s) * Cs[i] * l2;
        magS[i] = s;
      }
stresses:
  /* :254-274 wall level */
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      size_t i = (size_t)y * ld + x;
      double cst = Nu_t[i];
      txx[i] = -2.0 * cst * S[0][i];
      txy[i] = -2.0 * cst * S[1][i];
      tyy[i] = -2.0 * cst * S[3][i];
      tzz[i] = -2.0 * cst * S[5][i];
    }
  /* :288-325 */
#pragma omp parallel for
  for (int k = 1; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x, p = i + np;
        double cst = 0.5 * (Nu_t[i] + Nu_t[p]);
        txx[i] = -2.0 * cst * 0.5 * (S[0][i] + S[0][p]);
        txy[i] = -2.0 * cst * 0.5 * (S[1][i] + S[1][p]);
        tyy[i] = -2.0 * cst * 0.5 * (S[3][i] + S[3][p]);
        tzz[i] = -2.0 * cst * 0.5 * (S[5][i] + S[5][p]);
        double cn = Nu_t[i];
        txz[i] = -2.0 * cn * S[2][i];
        tyz[i] = -2.0 * cn * S[4][i];
      }
  /* :327-331 wall level txz, tyz were set by wallstress and are kept;
     :391-394 stress-free lid */
  for (size_t i = 0; i < np; ++i) { txz[(nz - 1) * np + i] = 0.0; tyz[(nz - 1) * np + i] = 0.0; }
  free(lz);
  return 0;
}

/* divstress_uv.f90:2-129 */
void or_divstress_uv(or_ctx *c, double *divt, const double *tx, const double *ty,
                     const double *tz) {
  const lp_grid *g = &c->g;
  size_t np = plane(c), nv = vol(c);
  int nz = g->nz, ld = g->ld;
  double *a = (double *)calloc(nv, sizeof(double)), *b = (double *)calloc(nv, sizeof(double));
  double *d = (double *)calloc(nv, sizeof(double));
  or_ddx(c, a, tx, nz);
  or_ddy(c, b, ty, nz);
  or_ddz_w(c, d, tz);
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < ld; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x;
        divt[i] = (x >= ld - 2) ? 0.0 : a[i] + b[i] + d[i];
      }
  for (size_t i = 0; i < np; ++i) divt[(nz - 1) * np + i] = BOGUS;
  free(a); free(b); free(d);
}

/* divstress_w.f90:4-85 */
void or_divstress_w(or_ctx *c, double *divt, const double *tx, const double *ty,
                    const double *tz) {
  const lp_grid *g = &c->g;
  size_t np = plane(c), nv = vol(c);
  int nz = g->nz, ld = g->ld;
  double *a = (double *)calloc(nv, sizeof(double)), *b = (double *)calloc(nv, sizeof(double));
  double *d = (double *)calloc(nv, sizeof(double));
  or_ddx(c, a, tx, nz);
  or_ddy(c, b, ty, nz);
  or_ddz_uv(c, d, tz);
  for (int k = 0; k < nz; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < ld; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x;
        double val;
        if (x >= ld - 2) val = 0.0;
        else if (k == 0) val = a[i] + b[i];
        else val = a[i] + b[i] + d[i];
        divt[i] = val;
      }
  free(a); free(b); free(d);
}

/* ------------------------------------------------------------------------- */
/* tridag_array.f90:1-163 -- Thomas algorithm per (kx,ky) mode, real          */
/* coefficients and complex right-hand side; rows j = 1..nz+1.               */
/* ------------------------------------------------------------------------- */
int or_tridag_array(or_ctx *c, const double *a, const double *b, const double *cc,
                    const double *r_, double *u_) {
  const lp_grid *g = &c->g;
  int lh = g->lh, ny = g->ny, n = g->nz + 1;
  size_t nh = (size_t)lh * ny;
  const cplx *r = (const cplx *)r_;
  cplx *u = (cplx *)u_;
  for (int jy = 0; jy < ny; ++jy)                                            /* :32-41 */
    for (int jx = 0; jx < lh - 1; ++jx)
      if (b[(size_t)jy * lh + jx] == 0.0) return fail(LP_ERR_PIVOT, "tridag_array: zero pivot in row 1");
  int err = 0;
#pragma omp parallel for
  for (size_t m = 0; m < nh; ++m) {
    int jx = (int)(m % lh), jy = (int)(m / lh);
    double bet = b[m];
    u[m] = r[m] / bet;                                                       /* :43-44 */
    if (jy == ny / 2 || jx == lh - 1 || (jx == 0 && jy == 0)) continue;      /* :61,65 */
    double *gam = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    for (int j = 1; j < n; ++j) {                                            /* :57-97 */
      size_t s = (size_t)j * nh + m;
      gam[j] = cc[s - nh] / bet;
      bet = b[s] - a[s] * gam[j];
      if (bet == 0.0) {
#pragma omp atomic write
        err = 1;
      }
      u[s] = (r[s] - a[s] * u[s - nh]) / bet;
    }
    for (int j = n - 2; j >= 0; --j) {                                       /* :125-155 */
      size_t s = (size_t)j * nh + m;
      u[s] = u[s] - gam[j + 1] * u[s + nh];
    }
    free(gam);
  }
  if (err) return fail(LP_ERR_PIVOT, "tridag_array: zero pivot");
  return 0;
}

/* ------------------------------------------------------------------------- */
/* press_stag_array.f90:1-361 (non-MPI; lbc = 1 so no surface-temperature    */
/* correction of the wall row)                                                */
/* ------------------------------------------------------------------------- */
int or_press_stag_array(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  int lh = g->lh, ny = g->ny, nz = g->nz, ld = g->ld;
  size_t np = plane(c), nh = (size_t)lh * ny;
  const double tadv1 = 1.5, dz = g->dz;
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W], *divtz = c->f[LP_DIVTZ];
  double *p = c->f[LP_P], *dpdx = c->f[LP_DPDX], *dpdy = c->f[LP_DPDY];
  cplx *H[3];
  for (int q = 0; q < 3; ++q) H[q] = (cplx *)calloc(nh * (nz + 1), sizeof(cplx)); /* levels 0..nz */
  const double cst = 1.0 / ((double)g->nx * g->ny);
  double *src[3] = {u, v, w};
#pragma omp parallel for schedule(dynamic)
  for (int kk = 0; kk < 3 * (nz - 1); ++kk) {                              /* :75-87 */
    int q = kk / (nz - 1), jz = kk % (nz - 1) + 1;
    const double *f = src[q] + (size_t)(jz - 1) * np;
    double *tmp = (double *)malloc(sizeof(double) * np);
    for (size_t i = 0; i < np; ++i) tmp[i] = cst / tadv1 * (f[i] / cf->dt);
    or_rfft2_fwd(c, 0, tmp, ld, (double *)(H[q] + (size_t)jz * nh));
    free(tmp);
  }
  /* :97-103: H_z(nz) = 0 */
  cplx *bottomw = (cplx *)malloc(sizeof(cplx) * nh), *topw = (cplx *)malloc(sizeof(cplx) * nh);
  {                                                                         /* :105-122 */
    double *tmp = (double *)malloc(sizeof(double) * np);
    for (size_t i = 0; i < np; ++i) tmp[i] = cst * divtz[i];
    or_rfft2_fwd(c, 0, tmp, ld, (double *)bottomw);
    for (size_t i = 0; i < np; ++i) tmp[i] = cst * divtz[(size_t)(nz - 1) * np + i];
    or_rfft2_fwd(c, 0, tmp, ld, (double *)topw);
    free(tmp);
  }
  for (int jz = 1; jz <= nz - 1; ++jz)                                      /* :126-131 */
    for (int q = 0; q < 3; ++q) kill_oddballs(H[q] + (size_t)jz * nh, lh, ny);
  kill_oddballs(topw, lh, ny);                                              /* :133-136 */
  kill_oddballs(bottomw, lh, ny);
  size_t n1 = nh * (nz + 1);
  double *a = (double *)calloc(n1, sizeof(double)), *b = (double *)calloc(n1, sizeof(double));
  double *cc = (double *)calloc(n1, sizeof(double));
  cplx *R = (cplx *)calloc(n1, sizeof(cplx));
  for (size_t m = 0; m < nh; ++m) {                                          /* :151-154, :174-177 */
    a[m] = BOGUS; b[m] = -1.0; cc[m] = 1.0; R[m] = -dz * bottomw[m];
    size_t t = (size_t)nz * nh + m;
    a[t] = -1.0; b[t] = 1.0; cc[t] = BOGUS; R[t] = -topw[m] * dz;
  }
  for (int jz = 2; jz <= nz; ++jz)                                           /* :229-258 */
    for (int jy = 0; jy < ny; ++jy) {
      if (jy == ny / 2) continue;
      for (int jx = 0; jx < lh - 1; ++jx) {
        if (jx == 0 && jy == 0) continue;
        size_t m = (size_t)jy * lh + jx, s = (size_t)(jz - 1) * nh + m;
        double kxv = c->kx[m], kyv = c->ky[m];
        a[s] = 1.0 / (dz * dz);
        b[s] = -(kxv * kxv + kyv * kyv + 2.0 / (dz * dz));
        cc[s] = 1.0 / (dz * dz);
        R[s] = I * (kxv * H[0][(size_t)(jz - 1) * nh + m] + kyv * H[1][(size_t)(jz - 1) * nh + m]) +
               (H[2][(size_t)jz * nh + m] - H[2][(size_t)(jz - 1) * nh + m]) / dz;
      }
    }
  cplx *ph = (cplx *)p;                                                       /* p_hat(lh,ny,0:nz) */
  int rc = or_tridag_array(c, a, b, cc, (const double *)R, (double *)ph);    /* :290 */
  if (rc == 0) {
    ph[0] = 0.0;                                                              /* :303-312 */
    ph[nh] = ph[0] - dz * bottomw[0];
    for (int jz = 2; jz <= nz; ++jz) ph[(size_t)jz * nh] = ph[(size_t)(jz - 1) * nh] + H[2][(size_t)jz * nh] * dz;
    for (int jz = 0; jz <= nz; ++jz) kill_oddballs(ph + (size_t)jz * nh, lh, ny); /* :328-329 */
    /* :339-352 */
    cplx *P0 = (cplx *)malloc(sizeof(cplx) * nh);
    memcpy(P0, ph, sizeof(cplx) * nh);
    or_rfft2_inv(c, 0, (double *)P0, p, ld);
    free(P0);
#pragma omp parallel for schedule(dynamic)
    for (int jz = 1; jz <= nz - 1; ++jz) {
      cplx *Ph = (cplx *)malloc(sizeof(cplx) * nh);
      cplx *Dx = (cplx *)malloc(sizeof(cplx) * nh), *Dy = (cplx *)malloc(sizeof(cplx) * nh);
      memcpy(Ph, ph + (size_t)jz * nh, sizeof(cplx) * nh);
      for (size_t m = 0; m < nh; ++m) {
        Dx[m] = I * c->kx[m] * Ph[m];
        Dy[m] = I * c->ky[m] * Ph[m];
      }
      or_rfft2_inv(c, 0, (double *)Dx, dpdx + (size_t)(jz - 1) * np, ld);
      or_rfft2_inv(c, 0, (double *)Dy, dpdy + (size_t)(jz - 1) * np, ld);
      or_rfft2_inv(c, 0, (double *)Ph, p + (size_t)jz * np, ld);
      free(Ph); free(Dx); free(Dy);
    }
    for (size_t i = 0; i < np; ++i) {                                          /* :355-357 */
      dpdx[(size_t)(nz - 1) * np + i] = BOGUS;
      dpdy[(size_t)(nz - 1) * np + i] = BOGUS;
      p[(size_t)nz * np + i] = BOGUS;
    }
  }
  for (int q = 0; q < 3; ++q) free(H[q]);
  free(bottomw); free(topw); free(a); free(b); free(cc); free(R);
  return rc;
}

/* forcing.f90:122-294 project (IBM force zero, inflow off, OCEAN off) */
int or_project(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld;
  size_t np = plane(c);
  const double tadv1 = 1.5;
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  double *dpdx = c->f[LP_DPDX], *dpdy = c->f[LP_DPDY], *dpdz = c->f[LP_DPDZ];
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x;
        double R = -tadv1 * dpdx[i];
        u[i] = (u[i] + cf->dt * (R + 0.0));
        R = -tadv1 * dpdy[i];
        v[i] = (v[i] + cf->dt * (R + 0.0));
        if (k >= 1) {
          R = -tadv1 * dpdz[i];
          w[i] = (w[i] + cf->dt * (R + 0.0));
        }
      }
  size_t top = (size_t)(nz - 1) * np, bel = (size_t)(nz - 2) * np;
  if (cf->use_mean_p_force)
    for (size_t i = 0; i < np; ++i) { u[top + i] = u[bel + i]; v[top + i] = v[bel + i]; }
  for (size_t i = 0; i < np; ++i) w[top + i] = 0.0;
  if (cf->lbc_mom == LP_LBC_STRESS_FREE)
    for (size_t i = 0; i < np; ++i) { u[i] = u[np + i]; v[i] = v[np + i]; }
  for (size_t i = 0; i < np; ++i) w[i] = 0.0;
  return 0;
}

/* scalars_module2.f90:1336-1379 */
void or_timestep_conditions(or_ctx *c, double *cfl, double *visc_stab) {
  const lp_grid *g = &c->g;
  size_t np = plane(c);
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W], *nu = c->f[LP_NU_T];
  double umax = 0.0, numax = -INFINITY;
  for (int k = 0; k < g->nz - 1; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * g->ld + x;
        double s = u[i] * u[i] + v[i] * v[i] + w[i] * w[i];
        if (s > umax) umax = s;
        if (nu[i] > numax) numax = nu[i];
      }
  double delta = fmin(g->dx, fmin(g->dy, g->dz));
  *cfl = sqrt(umax) * c->cfg.dt / delta;
  *visc_stab = c->cfg.dt * numax / (delta * delta);
}

/* rmsdiv.f90:1-44 */
void or_rmsdiv(or_ctx *c, double *rms) {
  const lp_grid *g = &c->g;
  size_t np = plane(c);
  double *du = c->f[LP_DUDX], *dv = c->f[LP_DVDY], *dw = c->f[LP_DWDZ];
  double s = 0.0;
  for (int k = 0; k < g->nz - 1; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * g->ld + x;
        s += fabs(du[i] + dv[i] + dw[i]);
      }
  *rms = s / ((double)g->nx * g->ny * (g->nz - 1));
}

/* io.f90:1782-1849 (avg_stats without MPI and theta): per level jz = 1..nz-1, row
 * k = jz-1 of out[(nz-1)][10] = <u>, <v>, <w>, <Cs_opt2>, <Nu_t>, <u'^2>, <v'^2>, <w'^2>,
 * <u'w'>, <txz>, plane means over x < nx.  u' and v' are moved onto the w-node (mean of
 * levels jz and jz-1, each minus its own plane mean), w' = w - <w>; at jz = 1 the
 * serial branch sets all three to zero (:1811-1815). */
/* energy.f90:1-47 (serial): ke = (u^2 + v^2 + w_uv^2)/2 on the uv nodes of levels
 * 1..nz-1, w_uv = (w(k) + w(k+1))/2, and its mean over those nx ny (nz-1) points, summed in
 * the reference's loop order (jz, jy, jx; each term divided by the count).  ke (volume,
 * may be null) receives the field. */
double or_energy(or_ctx *c, double *ke) {
  const lp_grid *g = &c->g;
  const size_t np = plane(c);
  const double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  const double denom = (double)g->nx * g->ny * (g->nz - 1);
  double kea = 0.0;
  for (int k = 0; k < g->nz - 1; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        const size_t i = (size_t)k * np + (size_t)y * g->ld + x;
        const double tw = 0.5 * (w[i] + w[i + np]);
        const double e = 0.5 * (u[i] * u[i] + v[i] * v[i] + tw * tw);
        if (ke) ke[i] = e;
        kea = kea + e / denom;
      }
  return kea;
}

/* io_crossflow_mod.f90:2647-2799 tavg_compute (serial, lbz = 1, levels 1..nz): acc gains
 * tavg_dt times each quantity, in the order of the accumulator struct:
 *   0 u, 1 v, 2 w_uv, 3 u_w, 4 v_w, 5 w, 6 u2, 7 v2, 8 w2, 9 uv, 10 uw, 11 vw,
 *   12 txx, 13 tyy, 14 tzz, 15 txy, 16 txz, 17 tyz, 18 p (p - (u^2 + w_uv^2 + v^2)/2),
 *   19 dudx, 20 dudy, 21 dudz_uv, 22 dvdx, 23 dvdy, 24 dvdz_uv, 25 dwdx_uv, 26 dwdy_uv,
 *   27 dwdz, 28 tot_pc, 29 tot_pc2, then per bin 30 + 2b pcon, 31 + 2b pcon2
 * (with droplets; the concentrations as volume fractions pcon pi/6 d^3).  Each quantity
 * is an (ld, ny, nz) volume; columns x >= nx are left alone.  interp_to_uv_grid
 * (functions.f90:53-128): f_uv(k) = (f(k) + f(k+1))/2, the top level repeats the one below;
 * interp_to_w_grid (:131-187): f_w(k) = (f(k-1) + f(k))/2, with u_w and v_w zeroed at the
 * bottom and top levels (:2688-2691).  The diffusive droplet fluxes nudPCondx/y are not
 * accumulated. */
void or_tavg_compute(or_ctx *c, double tavg_dt, double *acc) {
  const lp_grid *g = &c->g;
  const size_t np = plane(c), nv = vol(c);
  const int nz = g->nz, npc = c->cfg.drop.npcon;
  double **f = c->f;
  const double *u = f[LP_U], *v = f[LP_V], *w = f[LP_W], *p = f[LP_P];
  for (int k = 0; k < nz; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        const size_t i = (size_t)k * np + (size_t)y * g->ld + x;
        /* w-grid field to the uv grid */
        const int ku = k < nz - 1 ? k : nz - 2;
        const size_t iu = (size_t)ku * np + (size_t)y * g->ld + x;
#define UVG(F) (0.5 * (f[F][iu] + f[F][iu + np]))
        const double w_uv = 0.5 * (w[iu] + w[iu + np]);
        const double u_w = (k == 0 || k == nz - 1) ? 0.0 : 0.5 * (u[i - np] + u[i]);
        const double v_w = (k == 0 || k == nz - 1) ? 0.0 : 0.5 * (v[i - np] + v[i]);
        const double up = u[i], vp = v[i], wp = w[i];
        const double pres = p[i + np] - 0.5 * (up * up + w_uv * w_uv + vp * vp);
        double q[30];
        q[0] = up; q[1] = vp; q[2] = w_uv; q[3] = u_w; q[4] = v_w; q[5] = wp;
        q[6] = up * up; q[7] = vp * vp; q[8] = wp * wp; q[9] = up * vp; q[10] = u_w * wp; q[11] = v_w * wp;
        q[12] = f[LP_TXX][i]; q[13] = f[LP_TYY][i]; q[14] = f[LP_TZZ][i];
        q[15] = f[LP_TXY][i]; q[16] = f[LP_TXZ][i]; q[17] = f[LP_TYZ][i];
        q[18] = pres;
        q[19] = f[LP_DUDX][i]; q[20] = f[LP_DUDY][i]; q[21] = UVG(LP_DUDZ);
        q[22] = f[LP_DVDX][i]; q[23] = f[LP_DVDY][i]; q[24] = UVG(LP_DVDZ);
        q[25] = UVG(LP_DWDX); q[26] = UVG(LP_DWDY); q[27] = f[LP_DWDZ][i];
#undef UVG
        double tot = 0.0;
        for (int b = 0; b < npc; ++b) {
          const double d = c->cfg.drop.diameter[b];
          const double pc = f[LP_PCON][(size_t)b * nv + i] * PI_D / 6 * (d * d * d);
          tot = tot + pc;
          acc[(size_t)(30 + 2 * b) * nv + i] += pc * tavg_dt;
          acc[(size_t)(31 + 2 * b) * nv + i] += pc * pc * tavg_dt;
        }
        q[28] = tot; q[29] = tot * tot;
        const int nq = npc > 0 ? 30 : 28;
        for (int m = 0; m < nq; ++m) acc[(size_t)m * nv + i] += q[m] * tavg_dt;
      }
}

void or_plane_stats(or_ctx *c, double *out) {
  const lp_grid *g = &c->g;
  size_t np = plane(c);
  const double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  const double *cs = c->f[LP_CS_OPT2], *nu = c->f[LP_NU_T], *txz = c->f[LP_TXZ];
  const double nxy = (double)g->nx * g->ny;
  const int nl = g->nz - 1;
  for (int k = 0; k < nl; ++k) {
    double m[5] = {0, 0, 0, 0, 0};
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * g->ld + x;
        m[0] += u[i]; m[1] += v[i]; m[2] += w[i]; m[3] += cs[i]; m[4] += nu[i];
      }
    for (int q = 0; q < 5; ++q) out[k * 10 + q] = m[q] / nxy;
  }
  for (int k = 0; k < nl; ++k) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * g->ld + x;
        double up = 0.0, vp = 0.0, wp = 0.0;
        if (k > 0) {
          up = 0.5 * (u[i] - out[k * 10 + 0] + u[i - np] - out[(k - 1) * 10 + 0]);
          vp = 0.5 * (v[i] - out[k * 10 + 1] + v[i - np] - out[(k - 1) * 10 + 1]);
          wp = w[i] - out[k * 10 + 2];
        }
        s[0] += up * up; s[1] += vp * vp; s[2] += wp * wp; s[3] += up * wp; s[4] += txz[i];
      }
    for (int q = 0; q < 5; ++q) out[k * 10 + 5 + q] = s[q] / nxy;
  }
}

/* ------------------------------------------------------------------------- */
/* io.f90:1643-1679 (spectrum) on levels 1..nz-1 of a volume (bin `bin` of a    */
/* per-bin volume): each row u(1:nx, y)/nx is transformed forward; with X_m its */
/* m-th coefficient, E(m) = sum_y w_m |X_m|^2 / ny for m < nx/2, w_0 = 1/2 on   */
/* the real part, w_m = 1 otherwise; the Nyquist term is left out.  out[k*nh+m] */
/* ------------------------------------------------------------------------- */
int or_spectrum(or_ctx *c, int field, int bin, double *out) {
  const lp_grid *g = &c->g;
  if (field < 0 || field >= LP_NFIELDS || field == LP_P || field == LP_USTAR || bin < 0 ||
      (size_t)(bin + 1) * vol(c) > c->n[field])
    return 1;
  const int nh = g->nx / 2, nl = g->nz - 1;
  const double *f = c->f[field] + (size_t)bin * vol(c);
  fplan P;
  fplan_init(&P, g->nx);
  cplx *row = (cplx *)malloc(sizeof(cplx) * (size_t)g->nx);
  cplx *X = (cplx *)malloc(sizeof(cplx) * (size_t)g->nx);
  for (int k = 0; k < nl; ++k) {
    double *e = out + (size_t)k * nh;
    for (int m = 0; m < nh; ++m) e[m] = 0.0;
    for (int y = 0; y < g->ny; ++y) {
      const double *r = f + (size_t)k * plane(c) + (size_t)y * g->ld;
      for (int x = 0; x < g->nx; ++x) row[x] = r[x] / (double)g->nx;
      dft(&P, -1, X, row);
      e[0] += 0.5 * creal(X[0]) * creal(X[0]);
      for (int m = 1; m < nh; ++m) e[m] += creal(X[m]) * creal(X[m]) + cimag(X[m]) * cimag(X[m]);
    }
    for (int m = 0; m < nh; ++m) e[m] /= (double)g->ny;
  }
  free(row);
  free(X);
  fplan_free(&P);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* main.f90:864-898: ke = energy(); filt_da(ke); ddz_uv(ke); dudt += grad ke      */
/* (the acceleration the droplet inertia term reads at the next step)          */
/* ------------------------------------------------------------------------- */
void or_ke_dudt(or_ctx *c) {
  const lp_grid *g = &c->g;
  int nz = g->nz, ld = g->ld;
  size_t np = plane(c), nv = vol(c);
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  double *ke = (double *)calloc(nv, sizeof(double));
  for (int k = 0; k < nz - 1; ++k)                                              /* energy.f90:23-31 */
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x;
        double tw = 0.5 * (w[i] + w[i + np]);
        ke[i] = 0.5 * (u[i] * u[i] + v[i] * v[i] + tw * tw);
      }
  double *kt = (double *)malloc(nv * sizeof(double));
  memcpy(kt, ke, nv * sizeof(double));
  double *kx = (double *)calloc(nv, sizeof(double)), *ky = (double *)calloc(nv, sizeof(double));
  double *kz = (double *)calloc(nv, sizeof(double));
  or_filt_da(c, kt, kx, ky, nz);
  or_ddz_uv(c, kz, ke);
  for (size_t i = 0; i < (size_t)(nz - 1) * np; ++i) {
    c->f[LP_DUDT][i] += kx[i];
    c->f[LP_DVDT][i] += ky[i];
    c->f[LP_DWDT][i] += kz[i];
  }
  free(ke); free(kt); free(kx); free(ky); free(kz);
}

/* scalars_module_crossflow_mod.f90:4530-4597 calculate_probability: beta[i*n+j] for
 * a drop of bin j breaking into bin i (i < j) */
void or_calculate_probability(const double *d, int n, double sigma, double d_min, double *beta) {
  double *dc = (double *)malloc(sizeof(double) * n), *ds = (double *)malloc(sizeof(double) * n);
  double *emax = (double *)malloc(sizeof(double) * n), *emin = (double *)malloc(sizeof(double) * n);
  double *ef = (double *)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; ++i) { dc[i] = pow(d[i], 3.0); ds[i] = pow(d[i], 2.0); }
  for (int i = 0; i < n; ++i) {
    emax[i] = PI_D * sigma * 0.25992104 * ds[i];
    emin[i] = PI_D * sigma * (pow(dc[i] - d_min * d_min * d_min, 2.0 / 3.0) + d_min * d_min - ds[i]);
    for (int j = i + 1; j < n; ++j)
      ef[i * n + j] = PI_D * sigma * (pow(dc[j] - dc[i], 2.0 / 3.0) + ds[i] - ds[j]);
  }
  for (int i = 0; i < n * n; ++i) beta[i] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double num = emin[j] + emax[j] - ef[i * n + j];
      double den = 0.0;
      for (int k = 0; k < j; ++k) den += emin[j] + emax[j] - ef[k * n + j];
      beta[i * n + j] = num / den;
    }
  free(dc); free(ds); free(emax); free(emin); free(ef);
}

/* scalars_module_crossflow_mod.f90:2167-2230 breakage frequency of bin ip at a cell with
 * SGS dissipation `dissip` (fit parameters from the Ohnesorge number of the bin) */
static double breakage_freq(const lp_droplet_config *D, double dip, double dissip) {
  double Oh = pow(D->rho_c * D->nu_d * D->nu_d / D->sigma_s / dip, 0.5);
  double a = 2.392 * exp(-26.76 * Oh) + 2.877 * exp(-0.1244 * Oh);
  double b = 0.45;
  double cc = -pow(10.0, (0.5446 * pow(Oh, -0.3776) / (1 + 12.67 * pow(Oh, 1.462))));
  double dd = -(4.172 * pow(Oh, -0.5492)) / (1 + 0.5079 * pow(Oh, -0.4879));
  /* -0.7182 is a default-kind (binary32) literal in the reference */
  double e = pow(10.0, 0.4113 * exp(-55.94 * log10(Oh + 1)) + 0.5125 * exp(-(double)0.7182f * log10(Oh + 1)));
  double ds = dissip * pow(D->u_star, 3.0) / D->z_i;
  double tau_c = pow(fabs(dissip), 0.33333333) * pow(dip, -0.6666667);
  double Re = pow(fabs(ds), 1.0 / 3.0) * pow(dip, 4.0 / 3.0) / D->nu_c;
  double Re1 = log10(Re);
  double fb = 0.5 * ((signbit(Re1) ? -1.0 : 1.0) + 1.0);
  double al = a * fb * pow(Re1, fb * b);
  double bl = fb * cc * pow(Re1, fb * dd) - e;
  double fit = pow(10.0, al + bl);
  return D->k_b * pow(2.1, 0.5) * PI_D / 4.0 * 0.84 * tau_c * fit;
}

static inline double smart_r(double r) {
  return fmax(0.0, fmin(fmin(2.0 * r, 0.75 * r + 0.25), 4.0));
}
/* SMART face value given the upwind cell, the cell upwind of it, and the downwind
 * cell (:2877-3475): zero-denominator guard takes the upwind value */
static inline double smart_face(double cup, double cupup, double cdn) {
  double den = cup - cupup;
  if (den == 0.0) return cup;
  double r = smart_r((cdn - cup) / den);
  return cup + r * den / 2.0;
}

/* scalars_module_crossflow_mod.f90:1894-4472 pollen_RHS_calc2 (with the optional
 * breakup feedback into RHS_PCon).  PCon_scheme = 3 (SMART), model_sc = 1, lbcp = 1, no
 * Stokes drift, no SGS acceleration flux, single Gaussian point source. */
int or_pollen_rhs_calc2(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  const lp_droplet_config *D = &cf->drop;
  const int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld, npc = D->npcon;
  const size_t np = plane(c), nv = vol(c);
  const double dx = g->dx, dy = g->dy, dz = g->dz;
  double *PC = c->f[LP_PCON], *RH = c->f[LP_RHS_PCON], *PB = c->f[LP_RHS_PBE];
  double *u = c->f[LP_U], *v = c->f[LP_V], *w = c->f[LP_W];
  double *dudx = c->f[LP_DUDX], *dvdy = c->f[LP_DVDY], *nut = c->f[LP_NU_T], *magS = c->f[LP_MAGS];
  double *dudt = c->f[LP_DUDT], *dvdt = c->f[LP_DVDT], *dwdt = c->f[LP_DWDT];
  double *dissip = c->f[LP_DISSIP];
  if (npc < 1) return fail(LP_ERR_ARG, "pollen_RHS_calc2 needs npcon > 0");
  const int src_on = c->jt_total >= D->src_jt0 && c->jt_total < D->src_jt1;
  /* ---- pollen_RHS_calc2 ---- */
  /* face velocity arrays (ld, ny+1, nz) */
  size_t fp = (size_t)ld * (ny + 1);
#define FI(i, j, k) ((size_t)(k) * fp + (size_t)(j) * ld + (i))
  double *ui0 = (double *)calloc(fp * nz, sizeof(double)), *vi0 = (double *)calloc(fp * nz, sizeof(double));
  double *ui = (double *)calloc(fp * nz, sizeof(double)), *vi = (double *)calloc(fp * nz, sizeof(double));
  double *wi = (double *)calloc(fp * nz, sizeof(double));
  /* :2054-2081 u: dense nx x nx matrix times (dx dudx(1:nx-1), sum u) */
  double *dvec = (double *)malloc(sizeof(double) * (nx > ny ? nx : ny));
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const double *du = dudx + (size_t)k * np + (size_t)j * ld, *uu = u + (size_t)k * np + (size_t)j * ld;
      for (int m = 0; m < nx - 1; ++m) dvec[m] = dx * du[m];
      double su = 0.0;
      for (int m = 0; m < nx; ++m) su += uu[m];
      dvec[nx - 1] = su;
      for (int i = 0; i < nx; ++i) {
        double acc = 0.0;
        for (int m = 0; m < nx; ++m) {     /* matrix_x(m,i) */
          double M;
          if (m == nx - 1) M = 1.0 / nx;
          else if (i <= m) M = -(nx - (m + 1)) * (1.0 / nx);
          else M = (m + 1) * (1.0 / nx);
          acc += M * dvec[m];
        }
        ui0[FI(i, j, k)] = acc;
      }
      ui0[FI(nx, j, k)] = ui0[FI(0, j, k)];                                  /* :2081 */
    }
  /* :2088-2122 v */
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      for (int m = 0; m < ny - 1; ++m) dvec[m] = dy * dvdy[(size_t)k * np + (size_t)m * ld + i];
      double sv = 0.0;
      for (int m = 0; m < ny; ++m) sv += v[(size_t)k * np + (size_t)m * ld + i];
      dvec[ny - 1] = sv;
      for (int j = 0; j < ny; ++j) {
        double acc = 0.0;
        for (int m = 0; m < ny; ++m) {
          double M;
          if (m == ny - 1) M = 1.0 / ny;
          else if (j <= m) M = -(ny - (m + 1)) * (1.0 / ny);
          else M = (m + 1) * (1.0 / ny);
          acc += M * dvec[m];
        }
        vi0[FI(i, j, k)] = acc;
      }
      vi0[FI(i, ny, k)] = vi0[FI(i, 0, k)];
    }
  free(dvec);
  /* :2130 w is already on the faces; :2135-2138 magS(nz) = magS(nz-1) */
  for (size_t i = 0; i < np; ++i) magS[(size_t)(nz - 1) * np + i] = magS[(size_t)(nz - 2) * np + i];
  /* PBE set-up :2005-2021, :2167-2230 */
  double *beta = NULL, *gfreq = NULL;
  if (D->pbe) {
    memset(PB, 0, nv * npc * sizeof(double));
    beta = (double *)malloc(sizeof(double) * npc * npc);
    or_calculate_probability(D->diameter, npc, D->sigma_s, D->d_min, beta);
    gfreq = (double *)calloc(nv * npc, sizeof(double));
#pragma omp parallel for
    for (long q = 0; q < (long)nv * npc; ++q) {
      int ip = (int)(q / nv);
      size_t i = (size_t)(q % nv);
      gfreq[q] = breakage_freq(D, D->diameter[ip], dissip[i]);
    }
  }
  /* Kc_t on cell centres (:3482-3496) */
  double *Kc = (double *)calloc(nv, sizeof(double));
  for (int k = 0; k < nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t s = (size_t)k * np + (size_t)j * ld + i;
        double kv;
        if (k == 0) kv = nut[s];
        else if (k == nz - 2) kv = nut[s];
        else kv = (nut[s] + nut[s + np]) / 2.0;
        Kc[s] = kv / D->Sc;
      }
  const double ghat = D->g / (D->u_star * D->u_star / D->z_i);
  double *cx_ = (double *)malloc(sizeof(double) * fp * nz), *cy_ = (double *)malloc(sizeof(double) * fp * nz);
  double *cz_ = (double *)malloc(sizeof(double) * fp * nz);
  double *gx0 = (double *)malloc(sizeof(double) * ny * nz), *gxL = (double *)malloc(sizeof(double) * ny * nz);
  double *gy0 = (double *)malloc(sizeof(double) * nx * nz), *gyL = (double *)malloc(sizeof(double) * nx * nz);
  /* source footprint (:4222-4256): sum of weights; the 2 pi (width/dx)^2 cancels */
  double *wsrc = NULL, wsum = 0.0;
  if (src_on) {
    wsrc = (double *)malloc(sizeof(double) * nx * ny);
    double width = D->src_sigma * dx;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        double rx = ((i + 1) - D->src_i) * dx, ry = ((j + 1) - D->src_j) * dy;
        double r2 = rx * rx + ry * ry;
        double wv = exp(-pow(r2, D->src_ngau / 2) / (2.0 * pow(width, D->src_ngau)));
        wsrc[j * nx + i] = wv;
        wsum += wv / (2.0 * PI_D * pow(width / dx, 2.0));
      }
  }
  for (int ipc = 0; ipc < npc; ++ipc) {
    double *C = PC + (size_t)ipc * nv, *R = RH + (size_t)ipc * nv;
#define CC(i, j, k) C[(size_t)(k) * np + (size_t)(j) * ld + (i)]
    const double ws = D->settling_vel[ipc];
    /* :2243-2276 */
    memcpy(ui, ui0, sizeof(double) * fp * nz);
    memcpy(vi, vi0, sizeof(double) * fp * nz);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) wi[FI(i, j, k)] = w[(size_t)k * np + (size_t)j * ld + i];
    if (D->settling) {
      for (int k = 1; k < nz - 1; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) wi[FI(i, j, k)] -= ws;
      if (D->inertia) {
        double f = ws / ghat;
        for (int k = 0; k < nz; ++k)
          for (int j = 0; j < ny; ++j) {
            const double *a = dudt + (size_t)k * np + (size_t)j * ld;
            for (int i = 1; i < nx; ++i) ui[FI(i, j, k)] -= 0.5 * (a[i - 1] + a[i]) * f;
            ui[FI(0, j, k)] -= 0.5 * (a[nx - 1] + a[0]) * f;
            ui[FI(nx, j, k)] -= 0.5 * (a[nx - 1] + a[0]) * f;
          }
        for (int k = 0; k < nz; ++k)
          for (int i = 0; i < nx; ++i) {
            const double *a = dvdt + (size_t)k * np + i;
            for (int j = 1; j < ny; ++j) vi[FI(i, j, k)] -= 0.5 * (a[(size_t)(j - 1) * ld] + a[(size_t)j * ld]) * f;
            vi[FI(i, 0, k)] -= 0.5 * (a[(size_t)(ny - 1) * ld] + a[0]) * f;
            vi[FI(i, ny, k)] -= 0.5 * (a[(size_t)(ny - 1) * ld] + a[0]) * f;
          }
        for (int k = 1; k < nz - 1; ++k)
          for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) wi[FI(i, j, k)] -= dwdt[(size_t)k * np + (size_t)j * ld + i] * f;
      }
    }
    /* :2287-2467 surface flux, lbcp = 1 (pointsource: only negative fluxes kept) */
    double psf = D->sfc_flux;
    if (psf > 0) psf = 0.0;
    /* :2478-2590 ghosts and boundary faces */
    for (int k = 0; k < nz - 1; ++k) {
      for (int j = 0; j < ny; ++j) {
        size_t t = (size_t)k * ny + j;
        if (D->bc_x == 0) { gx0[t] = CC(nx - 1, j, k); gxL[t] = CC(0, j, k); }
        else if (D->bc_x == 1) {
          gxL[t] = CC(nx - 1, j, k); cx_[FI(0, j, k)] = 0.0; gx0[t] = 0.0; cx_[FI(nx, j, k)] = CC(nx - 1, j, k);
        } else {
          double s0 = signbit(ui[FI(0, j, k)]) ? -1.0 : 1.0;
          gx0[t] = CC(0, j, k) * (1 - s0) / 2; cx_[FI(0, j, k)] = gx0[t];
          double s1 = signbit(-ui[FI(0, j, k)]) ? -1.0 : 1.0;
          gxL[t] = CC(nx - 1, j, k) * (1 - s1) / 2; cx_[FI(nx, j, k)] = gxL[t];
        }
      }
      for (int i = 0; i < nx; ++i) {
        size_t t = (size_t)k * nx + i;
        if (D->bc_y == 0) { gy0[t] = CC(i, ny - 1, k); gyL[t] = CC(i, 0, k); }
        else if (D->bc_y == 1) {
          gy0[t] = CC(i, 0, k); cy_[FI(i, 0, k)] = CC(i, 0, k); cy_[FI(i, ny, k)] = 0.0; gyL[t] = 0.0;
        } else {
          double s0 = signbit(vi[FI(i, 0, k)]) ? -1.0 : 1.0;
          gy0[t] = CC(i, 0, k) * (1 - s0) / 2; cy_[FI(i, 0, k)] = gy0[t];
          double s1 = signbit(-vi[FI(i, 0, k)]) ? -1.0 : 1.0;
          gyL[t] = CC(i, ny - 1, k) * (1 - s1) / 2; cy_[FI(i, ny, k)] = gyL[t];
        }
      }
    }
    /* :2877-3475 SMART faces (face i lies between cells i-1 and i) */
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j) {
        size_t t = (size_t)k * ny + j;
        for (int i = 1; i < nx; ++i) {
          double cm2 = (i >= 2) ? CC(i - 2, j, k) : gx0[t];
          double cp1 = (i + 1 <= nx - 1) ? CC(i + 1, j, k) : gxL[t];
          if (ui[FI(i, j, k)] >= 0.0) cx_[FI(i, j, k)] = smart_face(CC(i - 1, j, k), cm2, CC(i, j, k));
          else cx_[FI(i, j, k)] = smart_face(CC(i, j, k), cp1, CC(i - 1, j, k));
        }
        if (D->bc_x == 0) {   /* :3370-3410 periodic faces 1 and nx+1 */
          double f0;
          if (ui[FI(0, j, k)] >= 0.0) f0 = smart_face(CC(nx - 1, j, k), CC(nx - 2, j, k), CC(0, j, k));
          else f0 = smart_face(CC(0, j, k), CC(1, j, k), CC(nx - 1, j, k));
          cx_[FI(0, j, k)] = f0;
          cx_[FI(nx, j, k)] = f0;
        }
      }
    for (int k = 0; k < nz - 1; ++k)
      for (int i = 0; i < nx; ++i) {
        size_t t = (size_t)k * nx + i;
        for (int j = 1; j < ny; ++j) {
          double cm2 = (j >= 2) ? CC(i, j - 2, k) : gy0[t];
          double cp1 = (j + 1 <= ny - 1) ? CC(i, j + 1, k) : gyL[t];
          if (vi[FI(i, j, k)] >= 0.0) cy_[FI(i, j, k)] = smart_face(CC(i, j - 1, k), cm2, CC(i, j, k));
          else cy_[FI(i, j, k)] = smart_face(CC(i, j, k), cp1, CC(i, j - 1, k));
        }
        if (D->bc_y == 0) {
          double f0;
          if (vi[FI(i, 0, k)] >= 0.0) f0 = smart_face(CC(i, ny - 1, k), CC(i, ny - 2, k), CC(i, 0, k));
          else f0 = smart_face(CC(i, 0, k), CC(i, 1, k), CC(i, ny - 1, k));
          cy_[FI(i, 0, k)] = f0;
          cy_[FI(i, ny, k)] = f0;
        }
      }
    /* z faces: face k (0-based) between cells k-1 and k */
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        for (int k = 2; k <= nz - 3; ++k) {
          if (wi[FI(i, j, k)] >= 0.0) cz_[FI(i, j, k)] = smart_face(CC(i, j, k - 1), CC(i, j, k - 2), CC(i, j, k));
          else cz_[FI(i, j, k)] = smart_face(CC(i, j, k), CC(i, j, k + 1), CC(i, j, k - 1));
        }
        /* face 2 (0-based 1): central when w >= 0 at the wall side */
        if (wi[FI(i, j, 1)] >= 0.0) cz_[FI(i, j, 1)] = (CC(i, j, 1) + CC(i, j, 0)) / 2.0;
        else cz_[FI(i, j, 1)] = smart_face(CC(i, j, 1), CC(i, j, 2), CC(i, j, 0));
        /* face nz-1 (0-based nz-2): uses C(nz) (= C(nz-1) after step_pollen2) */
        if (wi[FI(i, j, nz - 2)] >= 0.0)
          cz_[FI(i, j, nz - 2)] = smart_face(CC(i, j, nz - 3), CC(i, j, nz - 4), CC(i, j, nz - 2));
        else cz_[FI(i, j, nz - 2)] = smart_face(CC(i, j, nz - 2), CC(i, j, nz - 1), CC(i, j, nz - 3));
        cz_[FI(i, j, 0)] = 0.0;                                                /* :3340 */
        cz_[FI(i, j, nz - 1)] = cz_[FI(i, j, nz - 2)];                         /* :3356 */
      }
    /* :3559-3677 RHS = -div(u C) + div(Kc grad C) */
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t s = (size_t)k * np + (size_t)j * ld + i;
          double r = -(1.0 / dx) * (ui[FI(i + 1, j, k)] * cx_[FI(i + 1, j, k)] - ui[FI(i, j, k)] * cx_[FI(i, j, k)]);
          r = r - (1.0 / dy) * (vi[FI(i, j + 1, k)] * cy_[FI(i, j + 1, k)] - vi[FI(i, j, k)] * cy_[FI(i, j, k)]);
          r = r - (1.0 / dz) * (wi[FI(i, j, k + 1)] * cz_[FI(i, j, k + 1)] - wi[FI(i, j, k)] * cz_[FI(i, j, k)]);
          R[s] = r;
        }
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j) {
        size_t t = (size_t)k * ny + j;
        const double *Kr = Kc + (size_t)k * np + (size_t)j * ld;
        double *Rr = R + (size_t)k * np + (size_t)j * ld;
        const double *Cr = C + (size_t)k * np + (size_t)j * ld;
        for (int i = 1; i < nx - 1; ++i)
          Rr[i] = Rr[i] + (1.0 / (dx * dx)) * (((Kr[i] + Kr[i + 1]) / 2.0) * (Cr[i + 1] - Cr[i]) -
                                              ((Kr[i] + Kr[i - 1]) / 2.0) * (Cr[i] - Cr[i - 1]));
        Rr[0] = Rr[0] + (1.0 / (dx * dx)) * (((Kr[0] + Kr[1]) / 2.0) * (Cr[1] - Cr[0]) -
                                            ((Kr[0] + Kr[nx - 1]) / 2.0) * (Cr[0] - gx0[t]));
        Rr[nx - 1] = Rr[nx - 1] + (1.0 / (dx * dx)) * (((Kr[nx - 1] + Kr[0]) / 2.0) * (gxL[t] - Cr[nx - 1]) -
                                                      ((Kr[nx - 1] + Kr[nx - 2]) / 2.0) * (Cr[nx - 1] - Cr[nx - 2]));
      }
    for (int k = 0; k < nz - 1; ++k)
      for (int i = 0; i < nx; ++i) {
        size_t t = (size_t)k * nx + i;
#define KK(j) Kc[(size_t)k * np + (size_t)(j) * ld + i]
#define CY(j) C[(size_t)k * np + (size_t)(j) * ld + i]
#define RY(j) R[(size_t)k * np + (size_t)(j) * ld + i]
        for (int j = 1; j < ny - 1; ++j)
          RY(j) = RY(j) + (1.0 / (dy * dy)) * (((KK(j) + KK(j + 1)) / 2.0) * (CY(j + 1) - CY(j)) -
                                              ((KK(j) + KK(j - 1)) / 2.0) * (CY(j) - CY(j - 1)));
        RY(0) = RY(0) + (1.0 / (dy * dy)) * (((KK(0) + KK(1)) / 2.0) * (CY(1) - CY(0)) -
                                            ((KK(0) + KK(ny - 1)) / 2.0) * (CY(0) - gy0[t]));
        RY(ny - 1) = RY(ny - 1) + (1.0 / (dy * dy)) * (((KK(ny - 1) + KK(0)) / 2.0) * (gyL[t] - CY(ny - 1)) -
                                                      ((KK(ny - 1) + KK(ny - 2)) / 2.0) * (CY(ny - 1) - CY(ny - 2)));
#undef KK
#undef CY
#undef RY
      }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
#define KZ(k) Kc[(size_t)(k) * np + (size_t)j * ld + i]
#define CZ(k) C[(size_t)(k) * np + (size_t)j * ld + i]
#define RZ(k) R[(size_t)(k) * np + (size_t)j * ld + i]
        for (int k = 1; k < nz - 2; ++k)
          RZ(k) = RZ(k) + (1.0 / (dz * dz)) * (((KZ(k + 1) + KZ(k)) / 2.0) * (CZ(k + 1) - CZ(k)) -
                                              ((KZ(k) + KZ(k - 1)) / 2.0) * (CZ(k) - CZ(k - 1)));
        RZ(0) = RZ(0) + (1.0 / (dz * dz)) * (((KZ(1) + KZ(0)) / 2.0) * (CZ(1) - CZ(0))) + (1.0 / dz) * (psf);
        RZ(nz - 2) = RZ(nz - 2) + (1.0 / (dz * dz)) * (0.0 - ((KZ(nz - 2) + KZ(nz - 3)) / 2.0) * (CZ(nz - 2) - CZ(nz - 3)));
#undef KZ
#undef CZ
#undef RZ
      }
    /* :4222-4256 point source over two levels */
    if (src_on) {
      double width = D->src_sigma * dx;
      double q = D->Q_src[ipc];
      int k0 = D->src_k - 1;
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          double fw = wsrc[j * nx + i] / (2.0 * PI_D * pow(width / dx, 2.0)) / wsum;
          R[(size_t)k0 * np + (size_t)j * ld + i] += D->src_frac * q * fw;
          R[(size_t)(k0 + 1) * np + (size_t)j * ld + i] += (1.0 - D->src_frac) * q * fw;
        }
    }
    /* :4356-4448 breakup transfer from every larger bin */
    if (D->pbe) {
      for (int jp = ipc + 1; jp < npc; ++jp) {
        double bij = beta[ipc * npc + jp];
        const double di = D->diameter[ipc], dj = D->diameter[jp];
        double dc_i = di * di * di, dc_j = dj * dj * dj;                      /* diameter**3 */
        double d0 = pow(dc_j - dc_i, 1.0 / 3.0);
        int cnt = -1;
        double wlo = 0.0, whi = 0.0;
        if (d0 > D->diameter[0]) {
          for (int cn = 1; cn < npc; ++cn)
            if (!(d0 > D->diameter[cn])) {
              double a1 = fabs(D->diameter[cn] - d0), b1 = fabs(D->diameter[cn - 1] - d0);
              double ab = a1 + b1;
              cnt = cn;
              double dl = D->diameter[cn - 1], dh = D->diameter[cn];
              wlo = (a1 / ab) * ((d0 * d0 * d0) / (dl * dl * dl));
              whi = (b1 / ab) * ((d0 * d0 * d0) / (dh * dh * dh));
              break;
            }
        }
        for (int k = 0; k < nz - 1; ++k)
          for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
              size_t s = (size_t)k * np + (size_t)j * ld + i;
              double T = gfreq[(size_t)jp * nv + s] * bij * PC[(size_t)jp * nv + s];
              PB[(size_t)ipc * nv + s] += T;
              PB[(size_t)jp * nv + s] -= T;
              if (d0 > D->diameter[0]) {
                if (cnt > 0) {
                  PB[(size_t)(cnt - 1) * nv + s] += T * wlo;
                  PB[(size_t)cnt * nv + s] += T * whi;
                }
              } else {
                PB[s] += T * ((d0 * d0 * d0) / (D->diameter[0] * D->diameter[0] * D->diameter[0]));
              }
            }
      }
    }
#undef CC
  }
  /* :4470 (commented out in the reference): optional feedback of the breakup source */
  if (D->pbe && D->pbe_feedback)
    for (size_t i = 0; i < nv * npc; ++i) RH[i] += PB[i];
#undef FI
  free(ui0); free(vi0); free(ui); free(vi); free(wi); free(Kc);
  free(cx_); free(cy_); free(cz_); free(gx0); free(gxL); free(gy0); free(gyL);
  free(wsrc); free(beta); free(gfreq);
  return 0;
}

/* scalars_module_crossflow_mod.f90:1795-1843 calcbeta_pcon */
int or_calcbeta_pcon(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_droplet_config *D = &c->cfg.drop;
  const int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld, npc = D->npcon;
  const size_t np = plane(c), nv = vol(c);
  double *PC = c->f[LP_PCON];
  if (npc < 1) return fail(LP_ERR_ARG, "calcbeta_pcon needs npcon > 0");
  double *bp = c->f[LP_BETA_PCON];
  memset(bp, 0, nv * sizeof(double));
  if (D->active) {
    double gh = D->g * (D->z_i / (D->u_star * D->u_star));
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t s = (size_t)k * np + (size_t)j * ld + i;
          double t = 0.0;
          for (int ip = 0; ip < npc; ++ip) {
            double dd = D->diameter[ip];
            double V = PI_D / 6 * (dd * dd * dd);                             /* param.f90:665 */
            t = t + gh * (D->rho_c / D->rho_w - 1.0) * V * PC[(size_t)ip * nv + s];
          }
          bp[s] = t;
        }
  }
  return 0;
}

/* scalars_module_crossflow_mod.f90:4481-4508 step_pollen2 */
int or_step_pollen2(or_ctx *c) {
  const lp_grid *g = &c->g;
  const int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld, npc = c->cfg.drop.npcon;
  const size_t np = plane(c), nv = vol(c);
  const double dt = c->cfg.dt;
  double *PC = c->f[LP_PCON], *RH = c->f[LP_RHS_PCON], *RF = c->f[LP_RHS_PCON_F];
  if (npc < 1) return fail(LP_ERR_ARG, "step_pollen2 needs npcon > 0");
  for (int ip = 0; ip < npc; ++ip) {
    double *C = PC + (size_t)ip * nv, *R = RH + (size_t)ip * nv, *F = RF + (size_t)ip * nv;
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t s = (size_t)k * np + (size_t)j * ld + i;
          C[s] = C[s] + dt * (1.5 * R[s] - 0.5 * F[s]);
        }
    memcpy(C + (size_t)(nz - 1) * np, C + (size_t)(nz - 2) * np, np * sizeof(double));
  }
  return 0;
}

/* scalars_module_crossflow_mod.f90:1494-1786 pollen_all_in_one2: history, RHS, Euler
 * start, buoyancy, step (the mass-balance diagnostics at :1714-1781 are output only) */
int or_pollen_all_in_one2(or_ctx *c) {
  const lp_droplet_config *D = &c->cfg.drop;
  const size_t nv = vol(c);
  const int npc = D->npcon;
  int rc;
  if (npc < 1) return fail(LP_ERR_ARG, "pollen_all_in_one2 needs npcon > 0");
  double *RH = c->f[LP_RHS_PCON], *RF = c->f[LP_RHS_PCON_F];
  memcpy(RF, RH, nv * npc * sizeof(double));                                  /* :1521 */
  if ((rc = or_pollen_rhs_calc2(c))) return rc;
  if ((c->jt_total == D->pcon_init || (c->jt_total == D->src_jt0)) && !D->init_pcon)  /* :1660-1665 */
    memcpy(RF, RH, nv * npc * sizeof(double));
  if ((rc = or_calcbeta_pcon(c))) return rc;                                   /* :1673-1675 */
  return or_step_pollen2(c);                                                   /* :1700 */
}

/* ------------------------------------------------------------------------- */
/* Potential temperature (S_FLAG): scalars_module_crossflow_mod.f90 for a     */
/* homogeneous surface with a prescribed flux (patch_flag = 1, num_patch = 1, */
/* lbc = 1, spatial_flux_flag off, ubc = 0, U_stokes = ucross = 0).           */
/* ------------------------------------------------------------------------- */
/* :1170-1403 obukhov.  Homogeneous surface: the plane means of |U(1)|,        */
/* sgs_t3(1) and theta(1) give one u*, one Obukhov length and one set of       */
/* stability functions; before scal_init (or with theta off) the neutral       */
/* functions (:1183-1193).  zo is the uniform z0.                             */
int or_obukhov(or_ctx *c) {
  const lp_config *cf = &c->cfg;
  const lp_scalar_config *S = &cf->scal;
  const lp_grid *g = &c->g;
  double **f = c->f;
  const size_t np = plane(c);
  const double nxy = (double)g->nx * g->ny;
  const double *u = f[LP_U], *v = f[LP_V], *th = f[LP_THETA];
  double *psi = f[LP_PSI_M], *ust = f[LP_USTAR], *t3 = f[LP_SGS_T3];
  double *ob = c->obuk;
  double su = 0.0, spsi;
  for (int y = 0; y < g->ny; ++y)
    for (int x = 0; x < g->nx; ++x) {
      const size_t i = (size_t)y * g->ld + x;
      su += sqrt(u[i] * u[i] + v[i] * v[i]);
    }
#define PSUM(a, out)                                                        \
  do {                                                                      \
    out = 0.0;                                                              \
    for (int y = 0; y < g->ny; ++y)                                         \
      for (int x = 0; x < g->nx; ++x) out += (a)[(size_t)y * g->ld + x];    \
  } while (0)
  if (!S->s_flag || c->jt < S->scal_init) {                                 /* :1183-1193 */
    if (!S->initsc) memset(psi, 0, np * sizeof(double));
    PSUM(psi, spsi);
    const double us = su / nxy * cf->vonk / (log(0.5 * g->dz / cf->z0) - spsi / nxy);
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) ust[(size_t)y * g->ld + x] = us;
    ob[0] = spsi / nxy; ob[1] = 0.0; ob[2] = 1.0; ob[3] = 1.0;
    ob[4] = us; ob[5] = 0.0; ob[6] = 0.0; ob[7] = 0.0;
    return 0;
  }
  if (c->jt == S->scal_init && !S->initsc)                                  /* :1195-1201 */
    for (size_t i = 0; i < np; ++i) t3[i] = S->wt_s / S->u_star / S->T_scale;
  const double g_ = S->g / (S->u_star * S->u_star / S->z_i);                /* :1202 */
  double slz = 0.0;                                                         /* :1205 */
  for (int q = 0; q < g->nx * g->ny; ++q) slz += log(cf->z0);
  const double zo_ = exp(slz / nxy);
  double wt_, th_;
  PSUM(t3, wt_);
  wt_ /= nxy;                                                               /* :1206 */
  PSUM(psi, spsi);
  const double us = su / nxy * cf->vonk / (log(0.5 * g->dz / zo_) - spsi / nxy); /* :1207-1208 */
  PSUM(th, th_);
  th_ /= nxy;                                                               /* :1209 */
#undef PSUM
  /* :1210-1218 (homogeneous): wt_avg, ustar_avg, theta_avg are the means; :1232-1272 */
  const double L = -(us * us * us) / (cf->vonk * g_ / th_ * wt_);
  double pm = 0.0, ph = 0.0, fm = 1.0, fh = 1.0;
  if (L < 0.0 && wt_ != 0.0) {
    const double x = pow(1.0 - 16.0 * g->dz / 2.0 / L, 0.25);
    pm = 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + PI_D / 2.0;
    ph = 2.0 * log((1.0 + x * x) / 2.0);
    fm = 1.0 / x;
    fh = 1.0 / (x * x);
  } else if (L > 0.0 && wt_ != 0.0) {                 /* :1255-1258 (jan_diurnal_run off) */
    const double zeta = 0.5 * g->dz / L;
    fm = 1.0 + 4.8 * zeta;
    fh = 1.0 + 7.8 * zeta;
    pm = -4.8 * zeta;
    ph = -7.8 * zeta;
  }
  for (int y = 0; y < g->ny; ++y)
    for (int x = 0; x < g->nx; ++x) {
      const size_t i = (size_t)y * g->ld + x;
      psi[i] = pm;
      ust[i] = us;
    }
  ob[0] = pm; ob[1] = ph; ob[2] = fm; ob[3] = fh; ob[4] = us; ob[5] = L; ob[6] = wt_; ob[7] = th_;
  return 0;
}

void or_obukhov_get(or_ctx *c, double *out) { memcpy(out, c->obuk, sizeof c->obuk); }

/* :361-473 calcbeta (serial: levels 2..nz-1 get a value, level 1 and nz stay zero) */
static void or_calcbeta(or_ctx *c) {
  const lp_scalar_config *S = &c->cfg.scal;
  const lp_grid *g = &c->g;
  const size_t np = plane(c);
  const int nz = g->nz;
  const double *th = c->f[LP_THETA];
  double *beta = c->f[LP_BETA_SCAL];
  double *bar = (double *)malloc(sizeof(double) * nz);
  const double g_hat = S->g * (S->z_i / (S->u_star * S->u_star));
  memset(beta, 0, vol(c) * sizeof(double));
  for (int k = 0; k < nz; ++k) {                                            /* :375-383 */
    double s = 0.0;
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) s += th[(size_t)k * np + (size_t)y * g->ld + x];
    bar[k] = s / (double)(g->nx * g->ny);
  }
  const double at = 2.e-4 * S->T_scale;
  for (int k = 1; k < nz - 1; ++k)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        const size_t i = (size_t)k * np + (size_t)y * g->ld + x;
        const double a = th[i], b = th[i - np];
        double above, below, r;
        if (S->ocean && S->iboussinesq == 1) {                              /* :393-401 */
          above = 1.0 - at * a;
          below = 1.0 - at * b;
          r = -g_hat * (above + below) / 2.0;
        } else if (S->ocean && S->iboussinesq == 2) {                       /* :402-412 */
          above = at * (a - bar[k]) / (1.0 - at * bar[k]);
          below = at * (b - bar[k - 1]) / (1.0 - at * bar[k - 1]);
          r = g_hat * (above + below) / 2.0;
        } else if (S->ocean) {                                              /* :413-421 */
          above = at * a;
          below = at * b;
          r = g_hat * (above + below) / 2.0;
        } else {                                                            /* :427-435 */
          above = (a - bar[k]) / bar[k];
          below = (b - bar[k - 1]) / bar[k - 1];
          r = g_hat * (above + below) / 2.0;
        }
        beta[i] = r;
      }
  free(bar);
}

/* :141-360 theta_all_in_one with scalar_RHS_calc (:474-1110), calcbeta and
 * step_scalar (:1111-1169) */
int or_theta_all_in_one(or_ctx *c) {
  const lp_config *cf = &c->cfg;
  const lp_scalar_config *S = &cf->scal;
  const lp_grid *g = &c->g;
  double **f = c->f;
  const size_t np = plane(c), nv = vol(c);
  const int nz = g->nz, nx = g->nx, ny = g->ny, ld = g->ld;
  double *th = f[LP_THETA], *dTdx = f[LP_DTDX], *dTdy = f[LP_DTDY], *dTdz = f[LP_DTDZ];
  double *RHS = f[LP_RHS_T], *nut = f[LP_NU_T], *t3 = f[LP_SGS_T3];
  or_filt_da(c, th, dTdx, dTdy, nz);                                        /* :160 */
  or_ddz_uv(c, dTdz, th);                                                   /* :175 */
  for (size_t i = 0; i < np; ++i)                                           /* :176-178 */
    dTdz[(size_t)(nz - 1) * np + i] = S->dTdz_top / S->T_scale * S->z_i;
  memcpy(f[LP_RHS_TF], RHS, nv * sizeof(double));                           /* :199 */
  /* scalar_RHS_calc :530-575 (patch_flag = 1, one patch type): u* is the plane mean of
   * the wall law's ustar (a sum of ustar / patchnum); the flux wt_s passes through the default
   * REAL of surf_flux_current (:145, 489) */
  const double nxy = (double)(nx * ny);
  double ustar1 = 0.0;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) ustar1 += f[LP_USTAR][(size_t)y * ld + x] / nxy;
  const double wts = cf->fp32_casts ? (double)(float)S->wt_s : S->wt_s;
  const double sflux = wts / S->T_scale / S->u_star;
  const double phi_h = c->obuk[3];
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x)
      dTdz[(size_t)y * ld + x] = -phi_h * sflux / (ustar1 * cf->vonk * g->dz / 2.0);
  /* :808-813 dealias1 of u, v, w and the three derivatives; :827-874 products */
  const size_t npb = (size_t)g->ld_big * g->ny2, nvb = npb * nz;
  double *B[6];
  const double *src[6] = {f[LP_U], f[LP_V], f[LP_W], dTdx, dTdy, dTdz};
  for (int q = 0; q < 6; ++q) {
    B[q] = (double *)calloc(nvb, sizeof(double));
    or_dealias1(c, src[q], B[q], nz);
  }
  double *Rm = (double *)calloc(nvb, sizeof(double));
  for (int k = 0; k < nz; ++k)
    for (int y = 0; y < g->ny2; ++y)
      for (int x = 0; x < g->nx2; ++x) {
        const size_t i = (size_t)k * npb + (size_t)y * g->ld_big + x;
        double r = B[0][i] * B[3][i] + B[1][i] * B[4][i];
        if (k == 0)
          r = r + (0.5 * B[2][i + npb]) * B[5][i + npb];                    /* :852-853 */
        else if (k < nz - 1)
          r = r + (B[2][i] * B[5][i] + B[2][i + npb] * B[5][i + npb]) / 2.0; /* :837-838 */
        Rm[i] = r;                                                          /* :866 (top) */
      }
  or_dealias2(c, RHS, Rm, nz);                                              /* :879 */
  for (int q = 0; q < 6; ++q) free(B[q]);
  free(Rm);
  /* :889-1060 the SGS flux divergence; temp(nz) of the x and y fluxes is never set by the
   * reference (a local array), here zero: it reaches only RHS_T(nz), which no step reads */
  const double inv = 1.0 / S->Pr;
  double *tmp = (double *)calloc(nv, sizeof(double)), *dt_ = (double *)calloc(nv, sizeof(double));
  for (int dir = 0; dir < 2; ++dir) {
    const double *ds = dir == 0 ? dTdx : dTdy;
    memset(tmp, 0, nv * sizeof(double));
    for (int k = 0; k < nz - 1; ++k)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const size_t i = (size_t)k * np + (size_t)y * ld + x;
          tmp[i] = k == 0 ? inv * nut[i] * ds[i] : inv * 0.5 * (nut[i] + nut[i + np]) * ds[i];
        }
    if (dir == 0) {
      or_ddx(c, dt_, tmp, nz);                                              /* :925 */
      for (int k = 0; k < nz - 1; ++k)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)k * np + (size_t)y * ld + x;
            RHS[i] = -1.0 * RHS[i] + dt_[i];
          }
    } else {
      or_ddy(c, dt_, tmp, nz);                                              /* :977 */
      for (int k = 0; k < nz; ++k)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)k * np + (size_t)y * ld + x;
            RHS[i] = RHS[i] + dt_[i];
          }
    }
  }
  memset(tmp, 0, nv * sizeof(double));
  for (int k = 0; k < nz; ++k)                                              /* :990-1020 */
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t i = (size_t)k * np + (size_t)y * ld + x;
        if (k == 0) {
          tmp[i] = -1.0 * sflux;
          t3[i] = sflux;
        } else {
          tmp[i] = inv * nut[i] * dTdz[i];
          t3[i] = -1.0 * tmp[i];
        }
      }
  or_ddz_w(c, dt_, tmp);                                                    /* :1038 */
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t i = (size_t)k * np + (size_t)y * ld + x;
        RHS[i] = RHS[i] + dt_[i];
      }
  free(tmp);
  free(dt_);
  or_calcbeta(c);                                                           /* :246 */
  if (c->jt == S->scal_init && !S->initsc) memcpy(f[LP_RHS_TF], RHS, nv * sizeof(double)); /* :272-274 */
  const double *Rf = f[LP_RHS_TF];                                          /* :1111-1169 */
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t i = (size_t)k * np + (size_t)y * ld + x;
        th[i] = th[i] + cf->dt * (1.5 * RHS[i] - 0.5 * Rf[i]);
      }
  const double top = S->dTdz_top / S->T_scale * S->z_i * g->dz;
  for (size_t i = 0; i < np; ++i) {                                         /* :1155-1166 */
    const size_t it = (size_t)(nz - 1) * np + i;
    th[it] = g->L_z > 1.0 ? th[it - np] + top : th[it - np];
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* One time step: main.f90:207-826 (+:864-898 when droplets need dudt)       */
/* ------------------------------------------------------------------------- */
static int step_once(or_ctx *c) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  int nz = g->nz, nx = g->nx, ld = g->ld;
  size_t np = plane(c), nv = vol(c);
  double **f = c->f;
  int rc;
  memcpy(f[LP_RHSX_F], f[LP_RHSX], nv * sizeof(double));                   /* :252-254 */
  memcpy(f[LP_RHSY_F], f[LP_RHSY], nv * sizeof(double));
  memcpy(f[LP_RHSZ_F], f[LP_RHSZ], nv * sizeof(double));
  const int theta_on = cf->scal.s_flag && c->jt >= cf->scal.scal_init;
  if (cf->scal.s_flag) or_obukhov(c);                                       /* :257 */
  or_filt_da(c, f[LP_U], f[LP_DUDX], f[LP_DUDY], nz);                      /* :277-279 */
  or_filt_da(c, f[LP_V], f[LP_DVDX], f[LP_DVDY], nz);
  or_filt_da(c, f[LP_W], f[LP_DWDX], f[LP_DWDY], nz);
  or_ddz_uv(c, f[LP_DUDZ], f[LP_U]);                                        /* :284-288 */
  or_ddz_uv(c, f[LP_DVDZ], f[LP_V]);
  or_ddz_w(c, f[LP_DWDZ], f[LP_W]);
  if ((rc = or_wallstress(c))) return rc;                                   /* :315 */
  if ((rc = or_sgs_stag(c))) return rc;                                     /* :327 */
  if (theta_on) {                                                           /* :340-344 */
    if ((rc = or_theta_all_in_one(c))) return rc;
  } else if (cf->scal.s_flag) {
    memset(f[LP_BETA_SCAL], 0, nv * sizeof(double));
  }
  if (cf->drop.npcon > 0 && c->jt_total >= cf->drop.pcon_init) {                 /* :360-372 */
    if ((rc = or_pollen_all_in_one2(c))) return rc;
  } else {
    memset(f[LP_BETA_PCON], 0, nv * sizeof(double));                         /* :352 */
  }
  or_divstress_uv(c, f[LP_DIVTX], f[LP_TXX], f[LP_TXY], f[LP_TXZ]);         /* :394-397 */
  or_divstress_uv(c, f[LP_DIVTY], f[LP_TXY], f[LP_TYY], f[LP_TYZ]);
  or_divstress_w(c, f[LP_DIVTZ], f[LP_TXZ], f[LP_TYZ], f[LP_TZZ]);
  or_convec(c);                                                             /* :419 */
  for (size_t i = 0; i < (size_t)(nz - 1) * np; ++i) {                      /* :433-439 */
    f[LP_RHSX][i] = -f[LP_RHSX][i] - f[LP_DIVTX][i];
    f[LP_RHSY][i] = -f[LP_RHSY][i] - f[LP_DIVTY][i];
    f[LP_RHSZ][i] = -f[LP_RHSZ][i] - f[LP_DIVTZ][i];
    f[LP_DUDT][i] = -f[LP_DIVTX][i];
    f[LP_DVDT][i] = -f[LP_DIVTY][i];
    f[LP_DWDT][i] = -f[LP_DIVTZ][i];
    if (cf->scal.s_flag && !cf->scal.passive) f[LP_RHSZ][i] += f[LP_BETA_SCAL][i];   /* :450-453 */
    if (cf->drop.npcon > 0 && cf->drop.active) f[LP_RHSZ][i] += f[LP_BETA_PCON][i];  /* :458-460 */
  }
  double force = cf->use_mean_p_force ? cf->mean_p_force : 0.0;             /* :571-575 */
  if (c->jt == 1 && !cf->initu) {                                           /* :577-584 */
    memcpy(f[LP_RHSX_F], f[LP_RHSX], nv * sizeof(double));
    memcpy(f[LP_RHSY_F], f[LP_RHSY], nv * sizeof(double));
    memcpy(f[LP_RHSZ_F], f[LP_RHSZ], nv * sizeof(double));
  }
  const double tadv1 = 1.5, tadv2 = 1.0 - tadv1;
  for (size_t i = 0; i < (size_t)(nz - 1) * np; ++i) {                      /* :716-728 */
    f[LP_U][i] = f[LP_U][i] + cf->dt * (tadv1 * f[LP_RHSX][i] + tadv2 * f[LP_RHSX_F][i] + force);
    f[LP_V][i] = f[LP_V][i] + cf->dt * (tadv1 * f[LP_RHSY][i] + tadv2 * f[LP_RHSY_F][i]);
    f[LP_W][i] = f[LP_W][i] + cf->dt * (tadv1 * f[LP_RHSZ][i] + tadv2 * f[LP_RHSZ_F][i]);
    f[LP_DUDT][i] += cf->mean_p_force;
  }
  if ((rc = or_press_stag_array(c))) return rc;                             /* :785 */
  double *p = f[LP_P], *dpdz = f[LP_DPDZ];
  for (int k = 0; k < nz - 1; ++k)                                          /* :789-791 */
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < ld; ++x) {
        size_t i = (size_t)k * np + (size_t)y * ld + x;
        dpdz[i] = (x < nx) ? (p[i + np] - p[i]) / g->dz : 0.0;
      }
  for (size_t i = 0; i < np; ++i) dpdz[(size_t)(nz - 1) * np + i] = BOGUS;
  for (size_t i = 0; i < (size_t)(nz - 1) * np; ++i) {                      /* :795-803 */
    f[LP_RHSX][i] -= f[LP_DPDX][i];
    f[LP_RHSY][i] -= f[LP_DPDY][i];
    f[LP_RHSZ][i] -= f[LP_DPDZ][i];
    f[LP_DUDT][i] -= f[LP_DPDX][i];
    f[LP_DVDT][i] -= f[LP_DPDY][i];
    f[LP_DWDT][i] -= f[LP_DPDZ][i];
  }
  if ((rc = or_project(c))) return rc;                                      /* :826 */
  if (cf->drop.npcon > 0) or_ke_dudt(c);                                    /* :864-898 */
  return 0;
}

int or_step(or_ctx *c, int nsteps) {
  for (int s = 0; s < nsteps; ++s) {
    int rc = step_once(c);
    if (rc) return rc;
    c->jt += 1;
    c->jt_total += 1;
  }
  return 0;
}

/* ========================================================================= */
/* Discrete adjoint of the frozen-coefficient step (test infrastructure: the */
/* CPU restatement the HIP adjoint is checked against).                       */
/*                                                                           */
/* The frozen-coefficient step is step_once with Nu_t and the wall u* taken   */
/* from a base run (or_set_frozen) instead of the SGS model and the wall law: */
/* the map from (u, v, w, RHS_f) to (u, v, w, RHS) is then polynomial, and    */
/* or_step_adjoint applies the transpose of its linearisation at the state    */
/* held in the context.  Every transform below is the transpose of the        */
/* forward one: with the FFTW2 conventions (unnormalised, c2r using only the  */
/* real parts of the kx = 0 and kx = nx/2 terms) r2c^T = c2r after halving the */
/* interior kx columns, c2r^T = r2c followed by doubling them, the forward y  */
/* DFT's adjoint is the unnormalised inverse; the spectral derivatives with   */
/* the Nyquist row and column removed are antisymmetric, the filters and the  */
/* Nyquist removal symmetric; padding and truncation to and from the 3/2 grid */
/* are each other's transposes up to (nx2 ny2)/(nx ny).                       */
/* ========================================================================= */
void or_set_frozen(or_ctx *c, const double *nut, const double *ust) {
  c->frz_nut = nut;
  c->frz_ust = ust;
}

/* S <- W rfft2_fwd(f): the transpose of rfft2_inv (interior kx columns doubled) */
static void adj_c2r(or_ctx *c, const double *f, cplx *S) {
  const lp_grid *g = &c->g;
  or_rfft2_fwd(c, 0, f, g->ld, (double *)S);
  for (int y = 0; y < g->ny; ++y)
    for (int k = 1; k < g->lh - 1; ++k) S[(size_t)y * g->lh + k] *= 2.0;
}
/* f <- rfft2_inv(S / W): the transpose of rfft2_fwd (S is modified) */
static void adj_r2c(or_ctx *c, cplx *S, double *f) {
  const lp_grid *g = &c->g;
  for (int y = 0; y < g->ny; ++y)
    for (int k = 1; k < g->lh - 1; ++k) S[(size_t)y * g->lh + k] *= 0.5;
  or_rfft2_inv(c, 0, (const double *)S, f, g->ld);
}

/* d smart_face / d (cup, cupup, cdn) on the branch the values select (the face is linear
 * in the three cells on each branch of the limiter) */
static void smart_coef(double cup, double cupup, double cdn, double *a_up, double *a_upup, double *a_dn) {
  const double den = cup - cupup;
  *a_up = 1.0; *a_upup = 0.0; *a_dn = 0.0;
  if (den == 0.0) return;
  const double r = (cdn - cup) / den;
  const double b2 = 2.0 * r, bq = 0.75 * r + 0.25;
  const double m = fmin(fmin(b2, bq), 4.0);
  if (!(m > 0.0)) return;                                    /* psi = 0: the upwind cell */
  if (m == b2) { *a_up = 0.0; *a_dn = 1.0; return; }         /* psi = 2r: the downwind cell */
  if (m == bq) { *a_up = 0.75; *a_upup = -0.125; *a_dn = 0.375; return; }  /* QUICK */
  *a_up = 3.0; *a_upup = -2.0;                               /* psi = 4 */
}

/* Adjoint of the droplet part of one frozen-coefficient step (pollen_all_in_one2,
 * scalars_module_crossflow_mod.f90:1494-1786 with pollen_RHS_calc2 :1894-4472,
 * calcbeta_pcon :1795-1843 and step_pollen2 :4481-4508) at the base C (the context's
 * PCon) and the base filtered velocities uf, vf, wf with ux = dudx, vy = dvdy.  Frozen:
 * Nu_t (so Kc_t).  In: cb, rcb the adjoints of the step's PCon and RHS_PCon; arz the
 * adjoint of the assembled RHSz (which receives the buoyancy).  Out: cb, rcb those of the
 * PCon and RHS_PCon the step starts from; the velocity adjoints gain the transport's part
 * (auf, avf, awf of the filtered u, v, w; aux, avy of dudx, dvdy). */
static int adj_droplets(or_ctx *c, const double *nut, const double *uf, const double *vf, const double *wf,
                        const double *ux, const double *vy, const double *arz, double *cb, double *rcb,
                        double *auf, double *aux, double *avf, double *avy, double *awf) {
  const lp_grid *g = &c->g;
  const lp_droplet_config *D = &c->cfg.drop;
  const int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld, npc = D->npcon;
  const size_t np = plane(c), nv = vol(c);
  const double dx = g->dx, dy = g->dy, dz = g->dz, dt = c->cfg.dt;
  if (D->inertia) return fail(LP_ERR_ARG, "droplet adjoint: inertia (dudt in the face velocity) is not differentiated");
  if (D->pbe && D->pbe_feedback) return fail(LP_ERR_ARG, "droplet adjoint: the breakup feedback is not differentiated");
  if (c->jt_total < D->pcon_init) return 0;  /* no droplet step: PCon and RHS_PCon carry over */
  const int euler = (c->jt_total == D->pcon_init || c->jt_total == D->src_jt0) && !D->init_pcon;
  const double *PC = c->f[LP_PCON];
  const size_t fp = (size_t)ld * (ny + 1);
#define FI(i, j, k) ((size_t)(k) * fp + (size_t)(j) * ld + (i))
#define SI(i, j, k) ((size_t)(k) * np + (size_t)(j) * ld + (i))
  /* base face velocities (:2054-2138) */
  double *ui = (double *)calloc(fp * nz, sizeof(double)), *vi = (double *)calloc(fp * nz, sizeof(double));
  double *wi = (double *)calloc(fp * nz, sizeof(double));
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      double su = 0.0;
      for (int m = 0; m < nx; ++m) su += uf[SI(m, j, k)];
      for (int i = 0; i < nx; ++i) {
        double acc = 0.0;
        for (int m = 0; m < nx - 1; ++m)
          acc += (i <= m ? -(nx - (m + 1)) * (1.0 / nx) : (m + 1) * (1.0 / nx)) * dx * ux[SI(m, j, k)];
        ui[FI(i, j, k)] = acc + su / nx;
      }
      ui[FI(nx, j, k)] = ui[FI(0, j, k)];
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      double sv = 0.0;
      for (int m = 0; m < ny; ++m) sv += vf[SI(i, m, k)];
      for (int j = 0; j < ny; ++j) {
        double acc = 0.0;
        for (int m = 0; m < ny - 1; ++m)
          acc += (j <= m ? -(ny - (m + 1)) * (1.0 / ny) : (m + 1) * (1.0 / ny)) * dy * vy[SI(i, m, k)];
        vi[FI(i, j, k)] = acc + sv / ny;
      }
      vi[FI(i, ny, k)] = vi[FI(i, 0, k)];
    }
  double *Kc = (double *)calloc(nv, sizeof(double));
  for (int k = 0; k < nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t s0 = SI(i, j, k);
        double kv = (k == 0 || k == nz - 2) ? nut[s0] : 0.5 * (nut[s0] + nut[s0 + np]);
        Kc[s0] = kv / D->Sc;
      }
  /* adjoints of the face velocities, summed over the bins */
  double *aui = (double *)calloc(fp * nz, sizeof(double)), *avi = (double *)calloc(fp * nz, sizeof(double));
  double *awi = (double *)calloc(fp * nz, sizeof(double));
  double *mu = (double *)calloc(nv, sizeof(double)), *ac = (double *)calloc(nv, sizeof(double));
  double *acx = (double *)calloc(fp * nz, sizeof(double)), *acy = (double *)calloc(fp * nz, sizeof(double));
  double *acz = (double *)calloc(fp * nz, sizeof(double));
  const double gh = D->g * (D->z_i / (D->u_star * D->u_star));
  for (int ip = 0; ip < npc; ++ip) {
    const double *C = PC + (size_t)ip * nv;
    double *lb = cb + (size_t)ip * nv, *lr = rcb + (size_t)ip * nv;
    const double ws = D->settling_vel[ip];
#define CC(i, j, k) C[SI(i, j, k)]
    for (size_t q = 0; q < fp * nz; ++q) {
      wi[q] = 0.0; acx[q] = 0.0; acy[q] = 0.0; acz[q] = 0.0;
    }
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          wi[FI(i, j, k)] = wf[SI(i, j, k)];
          if (D->settling && k >= 1 && k <= nz - 2) wi[FI(i, j, k)] -= ws;
        }
    /* step_pollen2 (:4481-4508): top level copied from below; the update */
    memset(mu, 0, nv * sizeof(double));
    memset(ac, 0, nv * sizeof(double));
    for (size_t q = 0; q < np; ++q) {
      lb[(size_t)(nz - 2) * np + q] += lb[(size_t)(nz - 1) * np + q];
      lb[(size_t)(nz - 1) * np + q] = 0.0;
    }
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t s0 = SI(i, j, k);
          mu[s0] = lr[s0] + (euler ? dt : 1.5 * dt) * lb[s0];
          ac[s0] = lb[s0];                               /* C(k) carried by the update */
          lr[s0] = euler ? 0.0 : -0.5 * dt * lb[s0];     /* RHS_PCon_f = the RHS_PCon read */
        }
    for (size_t q = (size_t)(nz - 1) * np; q < nv; ++q) lr[q] = 0.0;
    /* calcbeta_pcon (:1795-1843) into RHSz (main.f90:458-460) */
    if (D->active) {
      const double dd = D->diameter[ip];
      const double coef = gh * (D->rho_c / D->rho_w - 1.0) * (PI_D / 6 * (dd * dd * dd));
      for (int k = 0; k < nz - 1; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) ac[SI(i, j, k)] += coef * arz[SI(i, j, k)];
    }
    /* RHS = -div(u C) (:3559-3594): adjoints of the face values and face velocities */
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
          double m = (i < nx ? mu[SI(i, j, k)] : 0.0) - (i >= 1 ? mu[SI(i - 1, j, k)] : 0.0);
          acx[FI(i, j, k)] += ui[FI(i, j, k)] / dx * m;
        }
    for (int k = 0; k < nz - 1; ++k)
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j <= ny; ++j) {
          double m = (j < ny ? mu[SI(i, j, k)] : 0.0) - (j >= 1 ? mu[SI(i, j - 1, k)] : 0.0);
          acy[FI(i, j, k)] += vi[FI(i, j, k)] / dy * m;
        }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int k = 0; k <= nz - 1; ++k) {
          double m = (k <= nz - 2 ? mu[SI(i, j, k)] : 0.0) - (k >= 1 ? mu[SI(i, j, k - 1)] : 0.0);
          acz[FI(i, j, k)] += wi[FI(i, j, k)] / dz * m;
        }
    /* the face values, recomputed, for the velocity adjoints; then the faces transposed */
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 0; j < ny; ++j) {
        /* x faces.  Ghosts (:2478-2590) */
        double gx0 = 0.0, gxL = 0.0, s0 = 0.0, s1 = 0.0;
        if (D->bc_x == 0) { gx0 = CC(nx - 1, j, k); gxL = CC(0, j, k); }
        else if (D->bc_x == 1) { gxL = CC(nx - 1, j, k); }
        else {
          s0 = signbit(ui[FI(0, j, k)]) ? 1.0 : 0.0;      /* (1 - sign)/2 */
          s1 = signbit(-ui[FI(0, j, k)]) ? 1.0 : 0.0;
          gx0 = CC(0, j, k) * s0; gxL = CC(nx - 1, j, k) * s1;
        }
        double agx0 = 0.0, agxL = 0.0;
        for (int i = 1; i < nx; ++i) {
          const size_t f = FI(i, j, k);
          const double m = (i < nx ? mu[SI(i, j, k)] : 0.0) - mu[SI(i - 1, j, k)];
          double cup, cupup, cdn, au, auu, ad;
          if (ui[f] >= 0.0) {
            cup = CC(i - 1, j, k); cupup = (i >= 2) ? CC(i - 2, j, k) : gx0; cdn = CC(i, j, k);
          } else {
            cup = CC(i, j, k); cupup = (i + 1 <= nx - 1) ? CC(i + 1, j, k) : gxL; cdn = CC(i - 1, j, k);
          }
          aui[f] += smart_face(cup, cupup, cdn) / dx * m;
          smart_coef(cup, cupup, cdn, &au, &auu, &ad);
          const double a = acx[f];
          if (ui[f] >= 0.0) {
            ac[SI(i - 1, j, k)] += au * a; ac[SI(i, j, k)] += ad * a;
            if (i >= 2) ac[SI(i - 2, j, k)] += auu * a; else agx0 += auu * a;
          } else {
            ac[SI(i, j, k)] += au * a; ac[SI(i - 1, j, k)] += ad * a;
            if (i + 1 <= nx - 1) ac[SI(i + 1, j, k)] += auu * a; else agxL += auu * a;
          }
        }
        {  /* boundary faces 0 and nx */
          const double m0 = mu[SI(0, j, k)], mL = -mu[SI(nx - 1, j, k)];
          const double a0 = acx[FI(0, j, k)], aL = acx[FI(nx, j, k)];
          if (D->bc_x == 0) {
            double cup, cupup, cdn, au, auu, ad;
            int iu, iuu, idn;
            if (ui[FI(0, j, k)] >= 0.0) { iu = nx - 1; iuu = nx - 2; idn = 0; }
            else { iu = 0; iuu = 1; idn = nx - 1; }
            cup = CC(iu, j, k); cupup = CC(iuu, j, k); cdn = CC(idn, j, k);
            const double f0 = smart_face(cup, cupup, cdn);
            aui[FI(0, j, k)] += f0 / dx * m0;
            aui[FI(nx, j, k)] += f0 / dx * mL;
            smart_coef(cup, cupup, cdn, &au, &auu, &ad);
            const double a = a0 + aL;
            ac[SI(iu, j, k)] += au * a; ac[SI(iuu, j, k)] += auu * a; ac[SI(idn, j, k)] += ad * a;
          } else if (D->bc_x == 1) {  /* cx(0) = 0, cx(nx) = C(nx-1) */
            aui[FI(nx, j, k)] += CC(nx - 1, j, k) / dx * mL;
            ac[SI(nx - 1, j, k)] += aL;
          } else {                    /* cx(0) = gx0, cx(nx) = gxL */
            aui[FI(0, j, k)] += gx0 / dx * m0;
            aui[FI(nx, j, k)] += gxL / dx * mL;
            agx0 += a0; agxL += aL;
          }
        }
        /* x diffusion (:3595-3622) with the ghosts */
        const double idx2 = 1.0 / (dx * dx);
        for (int i = 0; i < nx; ++i) {
          const double m = mu[SI(i, j, k)] * idx2;
          const double Ki = Kc[SI(i, j, k)];
          const double Kp = 0.5 * (Ki + Kc[SI((i + 1) % nx, j, k)]);
          const double Km = 0.5 * (Ki + Kc[SI((i + nx - 1) % nx, j, k)]);
          /* + Kp (C(i+1) - C(i)) - Km (C(i) - C(i-1)) */
          ac[SI(i, j, k)] -= (Kp + Km) * m;
          if (i + 1 <= nx - 1) ac[SI(i + 1, j, k)] += Kp * m; else agxL += Kp * m;
          if (i >= 1) ac[SI(i - 1, j, k)] += Km * m; else agx0 += Km * m;
        }
        /* ghosts to cells */
        if (D->bc_x == 0) { ac[SI(nx - 1, j, k)] += agx0; ac[SI(0, j, k)] += agxL; }
        else if (D->bc_x == 1) { ac[SI(nx - 1, j, k)] += agxL; }
        else { ac[SI(0, j, k)] += s0 * agx0; ac[SI(nx - 1, j, k)] += s1 * agxL; }
      }
    for (int k = 0; k < nz - 1; ++k)
      for (int i = 0; i < nx; ++i) {
        double gy0 = 0.0, gyL = 0.0, s0 = 0.0, s1 = 0.0;
        if (D->bc_y == 0) { gy0 = CC(i, ny - 1, k); gyL = CC(i, 0, k); }
        else if (D->bc_y == 1) { gy0 = CC(i, 0, k); }
        else {
          s0 = signbit(vi[FI(i, 0, k)]) ? 1.0 : 0.0;
          s1 = signbit(-vi[FI(i, 0, k)]) ? 1.0 : 0.0;
          gy0 = CC(i, 0, k) * s0; gyL = CC(i, ny - 1, k) * s1;
        }
        double agy0 = 0.0, agyL = 0.0;
        for (int j = 1; j < ny; ++j) {
          const size_t f = FI(i, j, k);
          const double m = mu[SI(i, j, k)] - mu[SI(i, j - 1, k)];
          double cup, cupup, cdn, au, auu, ad;
          if (vi[f] >= 0.0) {
            cup = CC(i, j - 1, k); cupup = (j >= 2) ? CC(i, j - 2, k) : gy0; cdn = CC(i, j, k);
          } else {
            cup = CC(i, j, k); cupup = (j + 1 <= ny - 1) ? CC(i, j + 1, k) : gyL; cdn = CC(i, j - 1, k);
          }
          avi[f] += smart_face(cup, cupup, cdn) / dy * m;
          smart_coef(cup, cupup, cdn, &au, &auu, &ad);
          const double a = acy[f];
          if (vi[f] >= 0.0) {
            ac[SI(i, j - 1, k)] += au * a; ac[SI(i, j, k)] += ad * a;
            if (j >= 2) ac[SI(i, j - 2, k)] += auu * a; else agy0 += auu * a;
          } else {
            ac[SI(i, j, k)] += au * a; ac[SI(i, j - 1, k)] += ad * a;
            if (j + 1 <= ny - 1) ac[SI(i, j + 1, k)] += auu * a; else agyL += auu * a;
          }
        }
        {
          const double m0 = mu[SI(i, 0, k)], mL = -mu[SI(i, ny - 1, k)];
          const double a0 = acy[FI(i, 0, k)], aL = acy[FI(i, ny, k)];
          if (D->bc_y == 0) {
            double au, auu, ad;
            int ju, juu, jdn;
            if (vi[FI(i, 0, k)] >= 0.0) { ju = ny - 1; juu = ny - 2; jdn = 0; }
            else { ju = 0; juu = 1; jdn = ny - 1; }
            const double cup = CC(i, ju, k), cupup = CC(i, juu, k), cdn = CC(i, jdn, k);
            const double f0 = smart_face(cup, cupup, cdn);
            avi[FI(i, 0, k)] += f0 / dy * m0;
            avi[FI(i, ny, k)] += f0 / dy * mL;
            smart_coef(cup, cupup, cdn, &au, &auu, &ad);
            const double a = a0 + aL;
            ac[SI(i, ju, k)] += au * a; ac[SI(i, juu, k)] += auu * a; ac[SI(i, jdn, k)] += ad * a;
          } else if (D->bc_y == 1) {  /* cy(0) = C(0), cy(ny) = 0 */
            avi[FI(i, 0, k)] += CC(i, 0, k) / dy * m0;
            ac[SI(i, 0, k)] += a0;
          } else {
            avi[FI(i, 0, k)] += gy0 / dy * m0;
            avi[FI(i, ny, k)] += gyL / dy * mL;
            agy0 += a0; agyL += aL;
          }
        }
        const double idy2 = 1.0 / (dy * dy);
        for (int j = 0; j < ny; ++j) {
          const double m = mu[SI(i, j, k)] * idy2;
          const double Kj = Kc[SI(i, j, k)];
          const double Kp = 0.5 * (Kj + Kc[SI(i, (j + 1) % ny, k)]);
          const double Km = 0.5 * (Kj + Kc[SI(i, (j + ny - 1) % ny, k)]);
          ac[SI(i, j, k)] -= (Kp + Km) * m;
          if (j + 1 <= ny - 1) ac[SI(i, j + 1, k)] += Kp * m; else agyL += Kp * m;
          if (j >= 1) ac[SI(i, j - 1, k)] += Km * m; else agy0 += Km * m;
        }
        if (D->bc_y == 0) { ac[SI(i, ny - 1, k)] += agy0; ac[SI(i, 0, k)] += agyL; }
        else if (D->bc_y == 1) { ac[SI(i, 0, k)] += agy0; }
        else { ac[SI(i, 0, k)] += s0 * agy0; ac[SI(i, ny - 1, k)] += s1 * agyL; }
      }
    /* z faces (:3285-3360) and z diffusion (:3623-3677) */
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        /* cz(nz-1) = cz(nz-2) */
        acz[FI(i, j, nz - 2)] += acz[FI(i, j, nz - 1)];
        for (int k = 1; k <= nz - 2; ++k) {
          const size_t f = FI(i, j, k);
          double cup, cupup, cdn, au, auu, ad, face;
          int ku, kuu, kdn;
          if (k == 1 && wi[f] >= 0.0) {  /* central at the wall side */
            face = 0.5 * (CC(i, j, 1) + CC(i, j, 0));
            ku = 1; kdn = 0; kuu = -1; au = 0.5; ad = 0.5; auu = 0.0;
          } else {
            if (wi[f] >= 0.0) { ku = k - 1; kuu = k - 2; kdn = k; }
            else { ku = k; kuu = k + 1; kdn = k - 1; }
            cup = CC(i, j, ku); cupup = CC(i, j, kuu); cdn = CC(i, j, kdn);
            face = smart_face(cup, cupup, cdn);
            smart_coef(cup, cupup, cdn, &au, &auu, &ad);
          }
          double m = mu[SI(i, j, k)] * (k <= nz - 2 ? 1.0 : 0.0) - mu[SI(i, j, k - 1)];
          double at = acz[f];
          if (k == nz - 2) {  /* face nz-1 holds the same value, carried by w(nz-1) */
            const double mt = -mu[SI(i, j, nz - 2)];
            awi[FI(i, j, nz - 1)] += face / dz * mt;
          }
          awi[f] += face / dz * m;
          ac[SI(i, j, ku)] += au * at; ac[SI(i, j, kdn)] += ad * at;
          if (kuu >= 0) ac[SI(i, j, kuu)] += auu * at;
        }
        const double idz2 = 1.0 / (dz * dz);
        for (int k = 0; k <= nz - 2; ++k) {
          const double m = mu[SI(i, j, k)] * idz2;
          if (k <= nz - 3) {  /* + Kp (C(k+1) - C(k)) */
            const double Kp = 0.5 * (Kc[SI(i, j, k + 1)] + Kc[SI(i, j, k)]);
            ac[SI(i, j, k + 1)] += Kp * m; ac[SI(i, j, k)] -= Kp * m;
          }
          if (k >= 1) {       /* - Km (C(k) - C(k-1)) */
            const double Km = 0.5 * (Kc[SI(i, j, k)] + Kc[SI(i, j, k - 1)]);
            ac[SI(i, j, k)] -= Km * m; ac[SI(i, j, k - 1)] += Km * m;
          }
        }
      }
    memcpy(lb, ac, nv * sizeof(double));
#undef CC
  }
  /* w faces: wi(k) = w(k) (less the settling velocity) */
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) awf[SI(i, j, k)] += awi[FI(i, j, k)];
  /* u faces: ui0 = M (dx dudx(0..nx-2), sum u), face nx = face 0 */
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      double l0 = aui[FI(0, j, k)] + aui[FI(nx, j, k)], lsum = 0.0;
      for (int i = 0; i < nx; ++i) lsum += (i == 0 ? l0 : aui[FI(i, j, k)]);
      for (int m = 0; m < nx - 1; ++m) {
        double acc = 0.0;
        for (int i = 0; i < nx; ++i)
          acc += (i <= m ? -(nx - (m + 1)) * (1.0 / nx) : (m + 1) * (1.0 / nx)) * (i == 0 ? l0 : aui[FI(i, j, k)]);
        aux[SI(m, j, k)] += dx * acc;
      }
      for (int m = 0; m < nx; ++m) auf[SI(m, j, k)] += lsum / nx;
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      double l0 = avi[FI(i, 0, k)] + avi[FI(i, ny, k)], lsum = 0.0;
      for (int j = 0; j < ny; ++j) lsum += (j == 0 ? l0 : avi[FI(i, j, k)]);
      for (int m = 0; m < ny - 1; ++m) {
        double acc = 0.0;
        for (int j = 0; j < ny; ++j)
          acc += (j <= m ? -(ny - (m + 1)) * (1.0 / ny) : (m + 1) * (1.0 / ny)) * (j == 0 ? l0 : avi[FI(i, j, k)]);
        avy[SI(i, m, k)] += dy * acc;
      }
      for (int m = 0; m < ny; ++m) avf[SI(i, m, k)] += lsum / ny;
    }
#undef FI
#undef SI
  free(ui); free(vi); free(wi); free(Kc); free(aui); free(avi); free(awi);
  free(mu); free(ac); free(acx); free(acy); free(acz);
  return 0;
}

/* adjoint of one frozen-coefficient step (main.f90:252-826 with the frozen Nu_t, u*).
 * Base: the context's u, v, w before the step and its jt (jt == 1 without initu: the
 * Euler start).  ub: 3 volumes, on entry the adjoint of the step's u, v, w, on return
 * that of the u, v, w it starts from; rb: 3 volumes, on entry the adjoint of the RHS the
 * step leaves, on return that of the RHS_f it reads.  The context's fields are used as
 * scratch (their values are undefined afterwards); droplets and the inflow are not
 * part of it. */
int or_step_adjoint(or_ctx *c, const double *nut, const double *ust, double *ub, double *rb) {
  return or_step_adjoint2(c, nut, ust, ub, rb, NULL, NULL);
}

/* or_step_adjoint with droplets: cb, rcb (npcon volumes each), on entry the adjoints of
 * the step's PCon and RHS_PCon, on return those of the PCon and RHS_PCon it starts from
 * (the base C is the context's PCon).  Required when npcon > 0. */
int or_step_adjoint2(or_ctx *c, const double *nut, const double *ust, double *ub, double *rb, double *cb,
                     double *rcb) {
  const lp_grid *g = &c->g;
  const lp_config *cf = &c->cfg;
  const int nx = g->nx, ny = g->ny, nz = g->nz, ld = g->ld, lh = g->lh;
  const size_t np = plane(c), nv = vol(c), nh = (size_t)lh * ny;
  const size_t npb = (size_t)g->ld_big * g->ny2;
  const double dz = g->dz, dt = cf->dt, cst = 1.0 / ((double)nx * ny);
  const int first = c->jt == 1 && !cf->initu;
  const int wall = cf->lbc_mom == LP_LBC_WALL;
  const double zl = 0.5 * dz - cf->d0, den = log(zl / cf->z0);
  if (cf->drop.npcon > 0) {  /* checked before anything is allocated */
    if (!cb || !rcb) return fail(LP_ERR_ARG, "step adjoint: droplets need the PCon and RHS_PCon adjoints");
    if (cf->drop.inertia) return fail(LP_ERR_ARG, "droplet adjoint: inertia (dudt in the face velocity) is not differentiated");
    if (cf->drop.pbe && cf->drop.pbe_feedback)
      return fail(LP_ERR_ARG, "droplet adjoint: the breakup feedback is not differentiated");
  }
#define NEWV(n) double *n = (double *)calloc(nv, sizeof(double))
#define X_ALL(k, y, x) for (int k = 0; k < nz; ++k) for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x)
#define IX(k, y, x) ((size_t)(k) * np + (size_t)(y) * ld + (x))
  /* ---- base intermediates of the step from u, v, w ---- */
  NEWV(uf); NEWV(vf); NEWV(wf);
  memcpy(uf, c->f[LP_U], nv * sizeof(double));
  memcpy(vf, c->f[LP_V], nv * sizeof(double));
  memcpy(wf, c->f[LP_W], nv * sizeof(double));
  NEWV(ux); NEWV(uy); NEWV(vx); NEWV(vy); NEWV(wx); NEWV(wy); NEWV(uz); NEWV(vz); NEWV(wz);
  or_filt_da(c, uf, ux, uy, nz);
  or_filt_da(c, vf, vx, vy, nz);
  or_filt_da(c, wf, wx, wy, nz);
  or_ddz_uv(c, uz, uf);
  or_ddz_uv(c, vz, vf);
  or_ddz_w(c, wz, wf);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      size_t i = (size_t)y * ld + x;
      uz[i] = wall ? uf[i] / (zl * den) : 0.0;
      vz[i] = wall ? vf[i] / (zl * den) : 0.0;
    }
  NEWV(om1); NEWV(om2); NEWV(om3);
  X_ALL(k, y, x) {
    size_t i = IX(k, y, x);
    if (k == 0) {
      om1[i] = 0.5 * (wy[i] + wy[i + np]) - vz[i];
      om2[i] = uz[i] - 0.5 * (wx[i] + wx[i + np]);
    } else {
      om1[i] = wy[i] - vz[i];
      om2[i] = uz[i] - wx[i];
    }
    om3[i] = vx[i] - uy[i];
  }
  double *U[3], *O[3], *aU[3], *aO[3];
  const double *srcU[3] = {uf, vf, wf}, *srcO[3] = {om1, om2, om3};
  for (int q = 0; q < 3; ++q) {
    U[q] = (double *)calloc(npb * nz, sizeof(double));
    O[q] = (double *)calloc(npb * nz, sizeof(double));
    aU[q] = (double *)calloc(npb * nz, sizeof(double));
    aO[q] = (double *)calloc(npb * nz, sizeof(double));
    or_dealias1(c, srcU[q], U[q], nz);
    or_dealias1(c, srcO[q], O[q], nz);
  }
  /* ---- adjoint ---- */
  double *au = ub, *av = ub + nv, *aw = ub + 2 * nv;
  double *rbx = rb, *rby = rb + nv, *rbz = rb + 2 * nv;
  NEWV(auf); NEWV(avf); NEWV(awf);
  /* forcing.f90:215-292 boundary values */
  for (size_t i = 0; i < np; ++i) {
    size_t top = (size_t)(nz - 1) * np + i, bel = (size_t)(nz - 2) * np + i;
    aw[i] = 0.0;
    aw[top] = 0.0;
    if (cf->lbc_mom == LP_LBC_STRESS_FREE) {
      au[np + i] += au[i]; au[i] = 0.0;
      av[np + i] += av[i]; av[i] = 0.0;
    }
    if (cf->use_mean_p_force) {
      au[bel] += au[top]; au[top] = 0.0;
      av[bel] += av[top]; av[top] = 0.0;
    } else {  /* the top level is filt_da's value */
      auf[top] += au[top]; au[top] = 0.0;
      avf[top] += av[top]; av[top] = 0.0;
    }
  }
  /* projection (forcing.f90:145-191) and RHS -= grad p (main.f90:795-803) */
  NEWV(aus); NEWV(avs); NEWV(aws); NEWV(apx); NEWV(apy); NEWV(apz); NEWV(arx); NEWV(ary); NEWV(arz);
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = IX(k, y, x);
        aus[i] = au[i]; apx[i] = -1.5 * dt * au[i] - rbx[i]; arx[i] = rbx[i];
        avs[i] = av[i]; apy[i] = -1.5 * dt * av[i] - rby[i]; ary[i] = rby[i];
        if (k >= 1) { aws[i] = aw[i]; apz[i] = -1.5 * dt * aw[i]; }
        apz[i] -= rbz[i];
        arz[i] = rbz[i];
      }
  /* dpdz = (p(k+1) - p(k))/dz (main.f90:789-791) */
  double *ap = (double *)calloc(np * (nz + 1), sizeof(double));
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = IX(k, y, x);
        ap[i + np] += apz[i] / dz;
        ap[i] -= apz[i] / dz;
      }
  /* p, dpdx, dpdy from p_hat (press_stag_array.f90:339-352) */
  cplx *aph = (cplx *)calloc(nh * (nz + 1), sizeof(cplx));
  {
    cplx *S = (cplx *)malloc(sizeof(cplx) * nh);
    for (int j = 0; j < nz; ++j) {
      adj_c2r(c, ap + (size_t)j * np, S);
      for (size_t m = 0; m < nh; ++m) aph[(size_t)j * nh + m] += S[m];
      if (j >= 1) {
        adj_c2r(c, apx + (size_t)(j - 1) * np, S);
        for (size_t m = 0; m < nh; ++m) aph[(size_t)j * nh + m] += -I * c->kx[m] * S[m];
        adj_c2r(c, apy + (size_t)(j - 1) * np, S);
        for (size_t m = 0; m < nh; ++m) aph[(size_t)j * nh + m] += -I * c->ky[m] * S[m];
      }
    }
    free(S);
  }
  /* the solve (press_stag_array.f90:149-329, tridag_array.f90): transposed systems */
  cplx *aH[3];
  for (int q = 0; q < 3; ++q) aH[q] = (cplx *)calloc(nh * (nz + 1), sizeof(cplx));
  cplx *abw = (cplx *)calloc(nh, sizeof(cplx)), *atw = (cplx *)calloc(nh, sizeof(cplx));
  {
    const double idz2 = 1.0 / (dz * dz);
    const int n = nz + 1;
    double *ga = (double *)malloc(sizeof(double) * n), *gb = (double *)malloc(sizeof(double) * n);
    double *gc = (double *)malloc(sizeof(double) * n), *gam = (double *)malloc(sizeof(double) * n);
    cplx *lam = (cplx *)malloc(sizeof(cplx) * n);
    for (size_t m = 0; m < nh; ++m) {
      int jx = (int)(m % lh), jy = (int)(m / lh);
      if (jy == ny / 2 || jx == lh - 1) continue;
      if (m == 0) { /* mean mode :303-312 */
        cplx acc = 0.0;
        for (int j = nz; j >= 2; --j) {
          acc += aph[(size_t)j * nh];
          if (j <= nz - 1) aH[2][(size_t)j * nh] += dz * acc;
        }
        acc += aph[nh];
        abw[0] += -dz * acc;
        continue;
      }
      double bi = -(c->kx[m] * c->kx[m] + c->ky[m] * c->ky[m] + 2.0 / (dz * dz));
      /* A: row 0 (b,c) = (-1, 1); rows 1..nz-1 (idz2, bi, idz2); row nz (a,b) = (-1, 1).
       * A^T: a'_j = c_{j-1}, b'_j = b_j, c'_j = a_{j+1} */
      for (int j = 0; j < n; ++j) {
        double aj = (j == 0) ? 0.0 : (j == nz ? -1.0 : idz2);
        double bj = (j == 0) ? -1.0 : (j == nz ? 1.0 : bi);
        double cj = (j == 0) ? 1.0 : (j == nz ? 0.0 : idz2);
        (void)aj; (void)cj;
        gb[j] = bj;
      }
      for (int j = 0; j < n; ++j) {
        double cprev = (j == 0) ? 0.0 : (j - 1 == 0 ? 1.0 : idz2);     /* c_{j-1} */
        double anext = (j == nz) ? 0.0 : (j + 1 == nz ? -1.0 : idz2);  /* a_{j+1} */
        ga[j] = cprev;
        gc[j] = anext;
      }
      double bet = gb[0];
      lam[0] = aph[m] / bet;
      for (int j = 1; j < n; ++j) {
        gam[j] = gc[j - 1] / bet;
        bet = gb[j] - ga[j] * gam[j];
        lam[j] = (aph[(size_t)j * nh + m] - ga[j] * lam[j - 1]) / bet;
      }
      for (int j = n - 2; j >= 0; --j) lam[j] -= gam[j + 1] * lam[j + 1];
      abw[m] += -dz * lam[0];
      atw[m] += -dz * lam[nz];
      for (int j = 1; j <= nz - 1; ++j) {
        aH[0][(size_t)j * nh + m] += -I * c->kx[m] * lam[j];
        aH[1][(size_t)j * nh + m] += -I * c->ky[m] * lam[j];
        aH[2][(size_t)j * nh + m] -= lam[j] / dz;
        if (j + 1 <= nz - 1) aH[2][(size_t)(j + 1) * nh + m] += lam[j] / dz;
      }
    }
    free(ga); free(gb); free(gc); free(gam); free(lam);
  }
  kill_oddballs(abw, lh, ny);
  kill_oddballs(atw, lh, ny);
  NEWV(adz);  /* adjoint of divtz */
  {
    cplx *S = (cplx *)malloc(sizeof(cplx) * nh);
    double *f = (double *)malloc(sizeof(double) * np);
    double *dst[3] = {aus, avs, aws};
    for (int q = 0; q < 3; ++q)
      for (int k = 0; k < nz - 1; ++k) {  /* H level k+1 from level k */
        memcpy(S, aH[q] + (size_t)(k + 1) * nh, sizeof(cplx) * nh);
        kill_oddballs(S, lh, ny);
        adj_r2c(c, S, f);
        for (size_t i = 0; i < np; ++i) dst[q][(size_t)k * np + i] += cst / (1.5 * dt) * f[i];
      }
    memcpy(S, abw, sizeof(cplx) * nh);
    adj_r2c(c, S, f);
    for (size_t i = 0; i < np; ++i) adz[i] += cst * f[i];
    memcpy(S, atw, sizeof(cplx) * nh);
    adj_r2c(c, S, f);
    for (size_t i = 0; i < np; ++i) adz[(size_t)(nz - 1) * np + i] += cst * f[i];
    free(S); free(f);
  }
  /* AB2 predictor (main.f90:716-728), RHS assembly (:433-439), Euler start (:577-584) */
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = IX(k, y, x);
        const double tr = first ? 1.0 : 1.5;
        auf[i] += aus[i]; avf[i] += avs[i]; awf[i] += aws[i];
        arx[i] += tr * dt * aus[i]; ary[i] += tr * dt * avs[i]; arz[i] += tr * dt * aws[i];
        rbx[i] = first ? 0.0 : -0.5 * dt * aus[i];
        rby[i] = first ? 0.0 : -0.5 * dt * avs[i];
        rbz[i] = first ? 0.0 : -0.5 * dt * aws[i];
      }
  for (int q = 0; q < 3; ++q)
    for (size_t i = (size_t)(nz - 1) * np; i < nv; ++i) rb[(size_t)q * nv + i] = 0.0;
  NEWV(acx); NEWV(acy); NEWV(acz); NEWV(adx); NEWV(ady);
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = IX(k, y, x);
        acx[i] = -arx[i]; adx[i] = -arx[i];
        acy[i] = -ary[i]; ady[i] = -ary[i];
        if (k >= 1) acz[i] = -arz[i];
        adz[i] += -arz[i];
      }
  /* convec.f90: products on the 3/2 grid (transposes of the padding and truncation) */
  NEWV(aux); NEWV(auy); NEWV(auz); NEWV(avx); NEWV(avy); NEWV(avz); NEWV(awx); NEWV(awy); NEWV(awz);
  {
    const double tt = ((double)nx * ny) / ((double)g->nx2 * g->ny2);  /* T^T = tt P */
    double *pb[3];
    const double *ac[3] = {acx, acy, acz};
    for (int q = 0; q < 3; ++q) {
      pb[q] = (double *)calloc(npb * nz, sizeof(double));
      or_dealias1(c, ac[q], pb[q], nz);
      for (size_t i = 0; i < npb * nz; ++i) pb[q][i] *= tt;
    }
    for (int k = 0; k < nz - 1; ++k)
      for (size_t i = 0; i < npb; ++i) {
        size_t a = (size_t)k * npb + i, b = a + npb, m = a - npb;
        double px = pb[0][a], py = pb[1][a], pz = pb[2][a];
        /* px = -U2 O3 + (U3(k+1) O2(k+1) + [k>=1] U3 O2)/2 */
        aU[1][a] += -O[2][a] * px; aO[2][a] += -U[1][a] * px;
        aU[2][b] += 0.5 * O[1][b] * px; aO[1][b] += 0.5 * U[2][b] * px;
        /* py = U1 O3 - (U3(k+1) O1(k+1) + [k>=1] U3 O1)/2 */
        aU[0][a] += O[2][a] * py; aO[2][a] += U[0][a] * py;
        aU[2][b] += -0.5 * O[0][b] * py; aO[0][b] += -0.5 * U[2][b] * py;
        if (k >= 1) {
          aU[2][a] += 0.5 * O[1][a] * px; aO[1][a] += 0.5 * U[2][a] * px;
          aU[2][a] += -0.5 * O[0][a] * py; aO[0][a] += -0.5 * U[2][a] * py;
          /* pz = (-(U1 + U1(k-1)) O2 + (U2 + U2(k-1)) O1)/2 */
          aU[0][a] += -0.5 * O[1][a] * pz; aU[0][m] += -0.5 * O[1][a] * pz;
          aO[1][a] += -0.5 * (U[0][a] + U[0][m]) * pz;
          aU[1][a] += 0.5 * O[0][a] * pz; aU[1][m] += 0.5 * O[0][a] * pz;
          aO[0][a] += 0.5 * (U[1][a] + U[1][m]) * pz;
        }
      }
    for (int q = 0; q < 3; ++q) free(pb[q]);
    const double pt = ((double)g->nx2 * g->ny2) / ((double)nx * ny);  /* P^T = pt T */
    double *aUf[3] = {auf, avf, awf};
    NEWV(tmp);
    NEWV(ao1); NEWV(ao2); NEWV(ao3);
    double *aom[3] = {ao1, ao2, ao3};
    for (int q = 0; q < 3; ++q) {
      or_dealias2(c, tmp, aU[q], nz);
      for (size_t i = 0; i < nv; ++i) aUf[q][i] += pt * tmp[i];
      or_dealias2(c, aom[q], aO[q], nz);
      for (size_t i = 0; i < nv; ++i) aom[q][i] *= pt;
    }
    free(tmp);
    /* vorticity (convec.f90:100-133; jz = 1 on the uvp node) */
    X_ALL(k, y, x) {
      size_t i = IX(k, y, x);
      if (k == 0) {
        awy[i] += 0.5 * ao1[i]; awy[i + np] += 0.5 * ao1[i];
        awx[i] -= 0.5 * ao2[i]; awx[i + np] -= 0.5 * ao2[i];
      } else {
        awy[i] += ao1[i];
        awx[i] -= ao2[i];
      }
      avz[i] -= ao1[i];
      auz[i] += ao2[i];
      avx[i] += ao3[i];
      auy[i] -= ao3[i];
    }
    free(ao1); free(ao2); free(ao3);
  }
  /* the droplet transport, its buoyancy and its update (main.f90:360-372) */
  if (cf->drop.npcon > 0) adj_droplets(c, nut, uf, vf, wf, ux, vy, arz, cb, rcb, auf, aux, avf, avy, awf);
  /* divstress_uv / divstress_w: the x, y derivatives are antisymmetric */
  NEWV(atxx); NEWV(atxy); NEWV(atxz); NEWV(atyy); NEWV(atyz); NEWV(atzz);
  {
    NEWV(t1);
    or_ddx(c, t1, adx, nz); for (size_t i = 0; i < nv; ++i) atxx[i] -= t1[i];
    or_ddy(c, t1, adx, nz); for (size_t i = 0; i < nv; ++i) atxy[i] -= t1[i];
    or_ddx(c, t1, ady, nz); for (size_t i = 0; i < nv; ++i) atxy[i] -= t1[i];
    or_ddy(c, t1, ady, nz); for (size_t i = 0; i < nv; ++i) atyy[i] -= t1[i];
    or_ddx(c, t1, adz, nz); for (size_t i = 0; i < nv; ++i) atxz[i] -= t1[i];
    or_ddy(c, t1, adz, nz); for (size_t i = 0; i < nv; ++i) atyz[i] -= t1[i];
    free(t1);
    for (int k = 0; k < nz - 1; ++k)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          size_t i = IX(k, y, x);
          atxz[i + np] += adx[i] / dz; atxz[i] -= adx[i] / dz;
          atyz[i + np] += ady[i] / dz; atyz[i] -= ady[i] / dz;
          if (k >= 1) { atzz[i] += adz[i] / dz; atzz[i - np] -= adz[i] / dz; }
        }
  }
  /* stresses with the frozen Nu_t (sgs_stag.f90:254-325) */
  NEWV(aS11); NEWV(aS12); NEWV(aS13); NEWV(aS22); NEWV(aS23); NEWV(aS33);
  X_ALL(k, y, x) {
    size_t i = IX(k, y, x);
    if (k == 0) {
      aS11[i] += -2.0 * nut[i] * atxx[i]; aS12[i] += -2.0 * nut[i] * atxy[i];
      aS22[i] += -2.0 * nut[i] * atyy[i]; aS33[i] += -2.0 * nut[i] * atzz[i];
    } else if (k <= nz - 2) {
      double cc = 0.5 * (nut[i] + nut[i + np]);
      aS11[i] += -cc * atxx[i]; aS11[i + np] += -cc * atxx[i];
      aS12[i] += -cc * atxy[i]; aS12[i + np] += -cc * atxy[i];
      aS22[i] += -cc * atyy[i]; aS22[i + np] += -cc * atyy[i];
      aS33[i] += -cc * atzz[i]; aS33[i + np] += -cc * atzz[i];
      aS13[i] += -2.0 * nut[i] * atxz[i];
      aS23[i] += -2.0 * nut[i] * atyz[i];
    }
  }
  /* strain rates (sgs_stag.f90:49-125) */
  X_ALL(k, y, x) {
    size_t i = IX(k, y, x), mm = i - np;
    if (k == 0) {
      aux[i] += aS11[i];
      auy[i] += 0.5 * aS12[i]; avx[i] += 0.5 * aS12[i];
      auz[i] += 0.5 * aS13[i]; awx[i] += 0.25 * aS13[i]; awx[i + np] += 0.25 * aS13[i];
      avy[i] += aS22[i];
      avz[i] += 0.5 * aS23[i]; awy[i] += 0.25 * aS23[i]; awy[i + np] += 0.25 * aS23[i];
      awz[i] += aS33[i];
    } else {
      aux[i] += 0.5 * aS11[i]; aux[mm] += 0.5 * aS11[i];
      auy[i] += 0.25 * aS12[i]; auy[mm] += 0.25 * aS12[i];
      avx[i] += 0.25 * aS12[i]; avx[mm] += 0.25 * aS12[i];
      auz[i] += 0.5 * aS13[i]; awx[i] += 0.5 * aS13[i];
      avy[i] += 0.5 * aS22[i]; avy[mm] += 0.5 * aS22[i];
      avz[i] += 0.5 * aS23[i]; awy[i] += 0.5 * aS23[i];
      awz[i] += 0.5 * aS33[i]; awz[mm] += 0.5 * aS33[i];
    }
  }
  /* z derivatives (ddz_uv, ddz_w) and the wall level (wallstress.f90 with frozen u*) */
  for (int k = 0; k < nz - 1; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = IX(k, y, x);
        if (k >= 1) {
          auf[i] += auz[i] / dz; auf[i - np] -= auz[i] / dz;
          avf[i] += avz[i] / dz; avf[i - np] -= avz[i] / dz;
        }
        awf[i + np] += awz[i] / dz; awf[i] -= awz[i] / dz;
      }
  if (wall) {
    double *qu = (double *)calloc(np, sizeof(double)), *qv = (double *)calloc(np, sizeof(double));
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t i = (size_t)y * ld + x;
        double q = -ust[i] * cf->vonk / den;
        auf[i] += auz[i] / (zl * den);
        avf[i] += avz[i] / (zl * den);
        qu[i] = q * atxz[i];
        qv[i] = q * atyz[i];
      }
    test_filter_plane(c, qu, c->G_test);
    test_filter_plane(c, qv, c->G_test);
    for (size_t i = 0; i < np; ++i) { auf[i] += qu[i]; avf[i] += qv[i]; }
    free(qu); free(qv);
  }
  /* filt_da: u -> (F u, Dx F u, Dy F u); transpose F(au) - Dx(aux) - Dy(auy) */
  {
    double *af[3] = {auf, avf, awf}, *ax[3] = {aux, avx, awx}, *ay[3] = {auy, avy, awy};
    NEWV(t1); NEWV(t2);
    for (int q = 0; q < 3; ++q) {
      double *out = ub + (size_t)q * nv;
      memcpy(out, af[q], nv * sizeof(double));
      or_filt_da(c, out, t1, t2, nz);
      or_ddx(c, t1, ax[q], nz);
      or_ddy(c, t2, ay[q], nz);
      for (size_t i = 0; i < nv; ++i) out[i] -= t1[i] + t2[i];
      for (int k = 0; k < nz; ++k)
        for (int y = 0; y < ny; ++y)
          for (int x = nx; x < ld; ++x) out[IX(k, y, x)] = 0.0;
    }
    free(t1); free(t2);
  }
  double *fr[] = {uf, vf, wf, ux, uy, vx, vy, wx, wy, uz, vz, wz, om1, om2, om3, auf, avf, awf, aus, avs, aws,
                  apx, apy, apz, arx, ary, arz, ap, adz, acx, acy, acz, adx, ady, aux, auy, auz, avx, avy, avz,
                  awx, awy, awz, atxx, atxy, atxz, atyy, atyz, atzz, aS11, aS12, aS13, aS22, aS23, aS33};
  for (size_t q = 0; q < sizeof fr / sizeof fr[0]; ++q) free(fr[q]);
  for (int q = 0; q < 3; ++q) { free(U[q]); free(O[q]); free(aU[q]); free(aO[q]); free(aH[q]); }
  free(aph); free(abw); free(atw);
#undef NEWV
#undef X_ALL
#undef IX
  return 0;
}
