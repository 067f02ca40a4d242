/*
 * lesgo_amd.h -- C ABI of the MI355X (gfx950) implementation of the per-timestep
 * RHS + pressure-projection path of the lesgo-lineage LES in
 * adiaiyer/Population_Balance.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an
 * int status (LP_OK == 0).  A Fortran host binds them through ISO_C_BINDING
 * (population_balance_amd/fortran/lesgo_amd_iface.f90); Python binds them through
 * ctypes (population_balance_amd/lesgo.py).
 *
 * Array conventions (identical to the reference's Fortran memory):
 *   - a 3-D field f(ld,ny,1:nz) is column-major, i.e. C f[k][jy][jx] with jx fastest,
 *     row pitch ld = 2*(nx/2+1), level k = jz-1.  Columns nx..ld-1 are FFT padding.
 *   - p(ld,ny,0:nz) has nz+1 levels, level k = jz.
 *   - "big" (3/2-padded) fields are (ld_big, ny2, nz).
 *   - spectra overlay the same memory as complex(lh,ny) (interleaved re/im).
 *   - element type: double for an LP_FP64 context, float for an LP_MIXED context
 *     (the exceptions, always double, are listed with the field ids below).  The
 *     explicit-array routines (lp_filt_da ... lp_divstress_w) take arrays of this type.
 *   - every array argument may be host memory or device memory: device pointers are
 *     used in place, host pointers are staged through the context (synchronous).  The
 *     context's streams do not wait on other streams: work that the caller queued on
 *     device arrays it passes must be complete before the call.  Calls return after
 *     their own work on the arrays has completed.
 *
 * Error behaviour: the reference STOPs on a zero pivot (tridag_array.f90:35-38,70-75)
 * and on an invalid lbc_mom (wallstress.f90:207-209); here those return
 * LP_ERR_PIVOT / LP_ERR_ARG and lp_last_error() holds the message.
 */
#ifndef LESGO_AMD_H
#define LESGO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LP_ABI_VERSION 2

enum lp_status {
  LP_OK = 0,
  LP_ERR_ARG = 1,         /* bad argument / unsupported configuration            */
  LP_ERR_HIP = 2,         /* HIP runtime error                                   */
  LP_ERR_PIVOT = 3,       /* zero pivot in the tridiagonal solve                 */
  LP_ERR_NAN = 4,         /* non-finite value: a velocity written by lp_step, or
                             any state field in lp_check_finite                 */
  LP_ERR_NOMEM = 5,       /* device allocation failed                            */
  LP_ERR_NODEVICE = 6,    /* no HIP device visible                               */
  LP_ERR_CFL = 7          /* CFL above the limit set by lp_set_cfl_limit         */
};

enum lp_precision {
  LP_FP64 = 0,            /* every field and every operation in binary64          */
  LP_MIXED = 1            /* FFT/convection/SGS in binary32; pressure tridiagonal,
                             mean mode, p, and the Lagrangian F_LM/F_MM/F_QN/F_NN
                             averages in binary64                                 */
};

enum lp_lbc_mom { LP_LBC_WALL = 0, LP_LBC_STRESS_FREE = 1 };

/* Grid: replaces the compile-time parameters of param.f90:57-62,70-72. */
typedef struct lp_grid {
  int32_t nx, ny, nz;            /* nz = staggered levels (reference nz, lbz = 1)      */
  int32_t ld, lh;                /* lh = nx/2+1, ld = 2*lh                             */
  int32_t nx2, ny2;              /* 3nx/2, 3ny/2                                       */
  int32_t lh_big, ld_big;        /* nx2/2+1, 2*lh_big                                  */
  int32_t pad_;
  double L_x, L_y, L_z;          /* domain; z_i = 1                                   */
  double dx, dy, dz;             /* L_x/nx, L_y/ny, L_z/(nz-1)                          */
} lp_grid;

/* Droplet (PCon) parameters: param.f90:303-440,560-682 re-expressed at run time.
 * Boundary modes for bc_x / bc_y: 0 periodic (periodicbcx/y), 1 inlet/outlet
 * (PCON_INLET_X: zero inflow at x = 0, zero gradient at x = L_x; PCON_INLET_Y: zero
 * gradient at y = 0, zero inflow at y = L_y), 2 upwind-sign ghosts (neither). */
typedef struct lp_droplet_config {
  int32_t npcon;                 /* number of size bins (0 = droplets off)             */
  int32_t pbe;                   /* compute the breakup source Rhs_pbe (PBE_FLAG)      */
  int32_t pbe_feedback;          /* add Rhs_pbe to RHS_PCon (commented out in the
                                    reference, scalars_module_crossflow_mod.f90:4470)  */
  int32_t settling, inertia;     /* settling, PCon_acc                                 */
  int32_t active;                /* active_pcon: droplet buoyancy into RHSz            */
  int32_t bc_x, bc_y;            /* see above                                          */
  int32_t src_i, src_j, src_k;   /* source centre xps, yps and level zps_local (1-based);
                                    src_frac goes to src_k, 1-src_frac to src_k+1      */
  int32_t src_ngau;              /* footprint exponent N_GAU (2 = Gaussian)            */
  int32_t src_jt0, src_jt1;      /* source on for src_jt0 <= jt_total < src_jt1       */
  int32_t pcon_init;             /* PCon_init: first jt_total with a droplet step      */
  int32_t init_pcon;             /* initPCon: RHS_PCon_f comes with the state          */
  double src_sigma;              /* footprint width in cells (2*b0_plume/dx)           */
  double src_frac;               /* share of the lower source level (0.7)              */
  double Sc;                     /* SGS Schmidt number (model_sc = 1)                  */
  double sfc_flux;               /* prescribed surface flux (lbcp = 1)                 */
  double rho_c, rho_w;           /* droplet / sea-water density                        */
  double sigma_s, nu_c, nu_d, k_b, d_min; /* surface tension, viscosities, breakup
                                    constant, minimum daughter size (param.f90:401)    */
  double g;                      /* gravity                                            */
  double u_star, z_i;            /* velocity and length scales                         */
  const double *diameter;        /* [npcon] bin diameters (m)                          */
  const double *settling_vel;    /* [npcon] rise/settling velocity (nondimensional)    */
  const double *Q_src;           /* [npcon] source strength (nondimensional)           */
} lp_droplet_config;

/* Potential temperature theta (S_FLAG): param.f90:244-283, 537, 616 and bottombc.f90:7
 * re-expressed at run time, for the reference's surface condition: a prescribed,
 * homogeneous surface flux (lbc = 1, patch_flag = 1, num_patch = 1).  lbc = 0 (surface
 * temperature through a log law) reads the scalar roughness z_os, which the reference's
 * sources never assign; lp_create refuses it.  The Stokes drift of the advection
 * (U_stokes) and the sponge of ubc = 1 are not modelled; the ocean's stress-free bottom
 * is lbc_mom = LP_LBC_STRESS_FREE. */
typedef struct lp_scalar_config {
  int32_t s_flag;                /* S_FLAG: theta on                                   */
  int32_t passive;               /* passive_scalar: no buoyancy beta_scal in RHSz       */
  int32_t scal_init;             /* SCAL_init: theta steps from jt >= scal_init        */
  int32_t initsc;                /* initsc: theta, RHS_T, sgs_t3(:,:,1), psi_m come with
                                    the state (no Euler start, psi_m kept)              */
  int32_t lbc;                   /* 1: prescribed surface flux wt_s                     */
  int32_t ocean;                 /* OCEAN_FLAG: calcbeta's ocean forms                  */
  int32_t iboussinesq;           /* 1, 2 or 3 with ocean (param.f90:616)                */
  int32_t pad_;
  double Pr;                     /* SGS Prandtl number (0.4)                            */
  double wt_s;                   /* surface flux, K m/s                                 */
  double T_scale;                /* 300 K                                               */
  double dTdz_top;               /* top gradient, K/m                                   */
  double g, u_star, z_i;         /* 9.81, 1, 1                                          */
} lp_scalar_config;

/* Run-time configuration: replaces param.f90 / sgsmodule.f90 parameters. */
typedef struct lp_config {
  lp_grid grid;
  double dt;                     /* param.f90:105                                      */
  int32_t precision;             /* enum lp_precision                                  */
  int32_t model;                 /* SGS model: 1 Smagorinsky, 5 Lagrangian scale-dep.  */
  int32_t inilag, initu;         /* param.f90:270                                      */
  int32_t cs_count, dyn_init;    /* param.f90:127,249                                  */
  int32_t fp32_casts;            /* reproduce REAL(4) conversions of lagrange_Sdep.f90:
                                    311,313,329,361-363,386,388,397,405-407,430-434    */
  int32_t lbc_mom;               /* enum lp_lbc_mom                                    */
  double z0, d0;                 /* roughness and displacement height (uniform)        */
  double vonk;                   /* 0.4                                                */
  double Co;                     /* model 1 constant (param.f90:160)                   */
  int32_t nnn;                   /* model 1 wall-damping exponent                      */
  int32_t use_mean_p_force;
  double mean_p_force;
  double cs2_init;               /* Cs_opt2 at jt == 1 with inilag (0.03)              */
  double opftime;                /* 1.5 (sgsmodule.f90:7)                              */
  double filter_size;            /* 1 (test_filtermodule.f90:8)                        */
  lp_droplet_config drop;
  lp_scalar_config scal;
} lp_config;

/* State field ids.  Shapes are (ld,ny,nz) unless noted.  In an LP_MIXED context the
 * fields marked [f64] stay binary64, all others are binary32.  The velocities are
 * binary64 in both: the binary32 transforms take u and v less a per-level value that is
 * added back in binary64, so the state keeps the mean flow without binary32 rounding. */
enum lp_field {
  LP_U = 0, LP_V, LP_W,          /* [f64]                                              */
  LP_DUDX, LP_DUDY, LP_DUDZ, LP_DVDX, LP_DVDY, LP_DVDZ, LP_DWDX, LP_DWDY, LP_DWDZ,
  LP_RHSX, LP_RHSY, LP_RHSZ, LP_RHSX_F, LP_RHSY_F, LP_RHSZ_F,
  LP_TXX, LP_TXY, LP_TXZ, LP_TYY, LP_TYZ, LP_TZZ,
  LP_DIVTX, LP_DIVTY, LP_DIVTZ,
  LP_P,                          /* (ld,ny,0:nz)  [f64]                                */
  LP_DPDX, LP_DPDY, LP_DPDZ,
  LP_F_LM, LP_F_MM, LP_F_QN, LP_F_NN,   /* [f64]                                      */
  LP_CS_OPT2, LP_BETA,
  LP_NU_T, LP_DISSIP, LP_MAGS,
  LP_U_LAG, LP_V_LAG, LP_W_LAG,
  LP_DUDT, LP_DVDT, LP_DWDT,     /* main.f90:433-439,795-803,864-898; lp_step forms them
                                    only when droplets are on (their only reader)       */
  LP_USTAR,                      /* (ld,ny) wall u* of wallstress                       */
  LP_PCON,                       /* (ld,ny,nz,npcon) [f64]                             */
  LP_RHS_PCON, LP_RHS_PCON_F,    /* (ld,ny,nz,npcon) [f64]                             */
  LP_RHS_PBE,                    /* (ld,ny,nz,npcon) [f64]                             */
  LP_BETA_PCON,                  /* (ld,ny,nz)                                         */
  LP_THETA,                      /* potential temperature / T_scale [f64]              */
  LP_RHS_T, LP_RHS_TF,           /* its RHS and the previous step's                    */
  LP_DTDX, LP_DTDY, LP_DTDZ,     /* dTdz: w nodes, level 1 from the surface flux        */
  LP_SGS_T3,                     /* vertical SGS flux (sgs_vert); level 1 = the surface
                                    flux [f64]                                          */
  LP_BETA_SCAL,                  /* theta buoyancy (calcbeta), added to RHSz            */
  LP_PSI_M,                      /* (ld,ny) stability correction of the wall law [f64] */
  LP_NFIELDS
};

typedef struct lp_ctx lp_ctx;

/* ---------------- lifecycle ---------------- */
int lp_abi_version(void);
const char *lp_last_error(void);
/* Fill the derived members of *g from nx, ny, nz and the domain (param.f90:57-72). */
int lp_grid_init(lp_grid *g, int32_t nx, int32_t ny, int32_t nz,
                 double L_x, double L_y, double L_z);
/* Defaults of the reference-equivalent configuration (model 5, inilag, cs_count=1,
 * DYN_init=2, wall BC, z0=1e-4, fp32_casts on, FP64). */
void lp_config_default(lp_config *c);
/* Create a context on the current HIP device: twiddles, wavenumbers (fft.f90:35-76),
 * G_test / G_test_test (test_filtermodule.f90:29-57, main.f90:179-183), device state,
 * scratch.  The droplet arrays of cfg->drop are copied. */
int lp_create(const lp_config *cfg, lp_ctx **out);
int lp_destroy(lp_ctx *ctx);
/* Wait for all work of the context.  The routines enqueue on the context's own HIP
 * streams; inside lp_step a second stream carries the convective term and the
 * single-plane passes and is joined into the first before the step ends.  Calls given
 * host arrays return after their copies back; with device arrays a call may return before
 * its kernels finish, so synchronise before the results are read on another stream.
 * lp_step returns when its steps have finished. */
int lp_sync(lp_ctx *ctx);

/* ---------------- state access ---------------- */
int lp_field_info(lp_ctx *ctx, int32_t field, void **dev_ptr, size_t *bytes,
                  int32_t *elem_size);
/* copy a whole field from / to `host` (host or device memory; synchronous) */
int lp_upload(lp_ctx *ctx, int32_t field, const void *host, size_t bytes);
int lp_download(lp_ctx *ctx, int32_t field, void *host, size_t bytes);
int lp_set_time(lp_ctx *ctx, int32_t jt, int32_t jt_total);
int lp_get_time(lp_ctx *ctx, int32_t *jt, int32_t *jt_total);

/* ---------------- spectral primitives (explicit arrays) ---------------- */
/* fft.f90:51-76 -> kx, ky, k2 (lh,ny) binary64 */
int lp_init_wavenumber(lp_ctx *ctx, double *kx, double *ky, double *k2);
/* test_filtermodule.f90:29-57 -> G (lh,ny) binary64, ifilter = 1 (sharp cutoff) */
int lp_test_filter_init(lp_ctx *ctx, double alpha, double *G);
/* filt_da.f90:3-36; overwrites f with its Nyquist-filtered self; nlev planes */
int lp_filt_da(lp_ctx *ctx, void *f, void *dfdx, void *dfdy, int32_t nlev);
/* ddx.f90:1-31, ddy.f90:1-31, ddxy.f90:1-29 */
int lp_ddx(lp_ctx *ctx, void *dfdx, const void *f, int32_t nlev);
int lp_ddy(lp_ctx *ctx, void *dfdy, const void *f, int32_t nlev);
int lp_ddxy(lp_ctx *ctx, void *dfdx, void *dfdy, const void *f, int32_t nlev);
/* ddz_uv.f90:5-60 (level 1 untouched), ddz_w.f90:6-51; nz levels */
int lp_ddz_uv(lp_ctx *ctx, void *dfdz, const void *f);
int lp_ddz_w(lp_ctx *ctx, void *dfdz, const void *f);
/* padd.f90:1-26 / unpadd.f90:1-22 on nplanes complex planes (lh,ny) <-> (lh_big,ny2) */
int lp_padd(lp_ctx *ctx, void *u_big, const void *u, int32_t nplanes);
int lp_unpadd(lp_ctx *ctx, void *cc, const void *cc_big, int32_t nplanes);
/* dealias1.f90:1-31 (u untouched) / dealias2.f90:1-33 (u_big is clobbered, as the
 * reference leaves its scaled spectrum there); nlev planes */
int lp_dealias1(lp_ctx *ctx, const void *u, void *u_big, int32_t nlev);
int lp_dealias2(lp_ctx *ctx, void *u, void *u_big, int32_t nlev);
/* test_filtermodule.f90:12-27 on nplanes (ld,ny) planes, G (lh,ny) binary64 */
int lp_test_filter(lp_ctx *ctx, void *f, const double *G, int32_t nplanes);

/* ---------------- physics routines (explicit arrays) ---------------- */
/* divstress_uv.f90:2-129 / divstress_w.f90:4-85 */
int lp_divstress_uv(lp_ctx *ctx, void *divt, const void *tx, const void *ty,
                    const void *tz);
int lp_divstress_w(lp_ctx *ctx, void *divt, const void *tx, const void *ty,
                   const void *tz);
/* tridag_array.f90:1-163: a,b,c (lh,ny,nz+1) binary64, r,u complex (lh,ny,nz+1)
 * binary64; modes jy = ny/2+1, jx = lh and (1,1) are skipped as in the reference. */
int lp_tridag_array(lp_ctx *ctx, const double *a, const double *b, const double *c,
                    const double *r, double *u);
/* No counterpart in the reference: the same solve for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No context; device arrays only, laid out as lp_tridag_array's with
 * (lh, ny, nz+1); enqueued on `stream` (a hipStream_t; null: the default stream) and
 * returns after the solve (LP_ERR_PIVOT on a zero pivot).  Its scratch -- the gam factors
 * and the pivot flag -- is one fixed device array of 72 MiB on each device the library
 * runs on, held until the process ends: lh*ny*(nz+1) <= 9*2^20 - 32 values (256^3 fits),
 * larger systems return LP_ERR_ARG.  Calls that overlap on one device race. */
int lp_tridag_array_noalloc(int32_t lh, int32_t ny, int32_t nz, const double *a, const double *b,
                            const double *c, const double *r, double *u, void *stream);

/* ---------------- physics routines on the context state ----------------
 * The reference's no-argument routines read and write module globals
 * (sim_param, sgsmodule); these read and write the context's fields. */
int lp_wallstress(lp_ctx *ctx);                 /* wallstress.f90:5-215            */
int lp_sgs_stag(lp_ctx *ctx);                   /* sgs_stag.f90:3-401 (+lagrange)  */
int lp_lagrange_sdep(lp_ctx *ctx);              /* lagrange_Sdep.f90:4-496         */
int lp_interpolag_sdep(lp_ctx *ctx);            /* interpolag_Sdep.f90:2-373       */
int lp_convec(lp_ctx *ctx);                     /* convec.f90:9-316 -> RHSx/y/z    */
int lp_press_stag_array(lp_ctx *ctx);           /* press_stag_array.f90:1-361      */
int lp_project(lp_ctx *ctx);                    /* forcing.f90:122-294             */
int lp_timestep_conditions(lp_ctx *ctx, double *cfl, double *visc_stab);
                                                /* scalars_module2.f90:1336-1379   */
int lp_rmsdiv(lp_ctx *ctx, double *rmsdiv);     /* rmsdiv.f90:1-44                 */
/* Potential temperature (cfg.scal.s_flag; LP_ERR_ARG without it).
 * lp_obukhov: scalars_module_crossflow_mod.f90:1170-1403 obukhov for a homogeneous
 * surface: from u, v, theta and sgs_t3 on level 1 and psi_m, the plane-mean u*, the
 * Monin-Obukhov length L and the stability functions; psi_m (LP_PSI_M, read by the wall
 * law) and the u* plane (LP_USTAR) are rewritten with them.  lp_step calls it at the
 * start of each step (main.f90:257) when theta is on.  lp_obukhov_get returns the last
 * call's values: out[0..7] = psi_m, psi_h, phi_m, phi_h, u*, L, the mean surface flux,
 * the mean theta of level 1 (L = 0 and the neutral functions before scal_init).
 * lp_theta_all_in_one: :141-360 theta_all_in_one (filt_da and ddz_uv of theta,
 * scalar_RHS_calc :474-1110 with the 3/2-rule advection and the SGS flux divergence,
 * calcbeta :361-473, step_scalar :1111-1169); lp_step calls it after sgs_stag
 * (main.f90:340-344) once jt >= scal_init and adds beta_scal to RHSz (:450-453). */
int lp_obukhov(lp_ctx *ctx);
int lp_obukhov_get(lp_ctx *ctx, double *out);
int lp_theta_all_in_one(lp_ctx *ctx);
/* io.f90:1782-1849 (avg_stats, serial case, without theta): the instantaneous plane
 * statistics that avg_stats time-averages and writes.  out (host, binary64) holds
 * (nz-1) rows of 10: <u>, <v>, <w>, <Cs_opt2>, <Nu_t>, <u'^2>, <v'^2>, <w'^2>, <u'w'>,
 * <txz> over x < nx; u', v' on the w-node, all three fluctuations zero on the first
 * level.  The running averages and the output files stay with the caller. */
int lp_plane_stats(lp_ctx *ctx, double *out);
/* energy.f90:1-47: kea = mean of (u^2 + v^2 + w_uv^2)/2 over the uv nodes of levels
 * 1..nz-1 (w_uv = (w(k) + w(k+1))/2), binary64. */
int lp_energy(lp_ctx *ctx, double *kea);
/* Running time averages, io_crossflow_mod.f90:2452-2799.  lp_tavg_reset is tavg_init (the
 * accumulators and tavg_total_time start at zero); lp_tavg_compute(ctx, tavg_dt) is
 * tavg_compute: every accumulator gains tavg_dt times its quantity at the current state
 * (the host counts tavg_dt as the reference's output loop does, :255-280).  Quantity q
 * (binary64 (ld,ny,nz), levels 1..nz): 0 u, 1 v, 2 w_uv, 3 u_w, 4 v_w, 5 w, 6 u2, 7 v2,
 * 8 w2, 9 uv, 10 uw, 11 vw, 12 txx, 13 tyy, 14 tzz, 15 txy, 16 txz, 17 tyz,
 * 18 p - (u^2 + w_uv^2 + v^2)/2, 19 dudx, 20 dudy, 21 dudz_uv, 22 dvdx, 23 dvdy, 24 dvdz_uv,
 * 25 dwdx_uv, 26 dwdy_uv, 27 dwdz; with droplets 28 tot_pc, 29 tot_pc2, 30+2b pcon and
 * 31+2b pcon2 of bin b (volume fractions pcon pi/6 d^3).  The droplets' diffusive fluxes
 * nudPCondx/y are not accumulated.  lp_tavg_get copies quantity q (host or device out)
 * and returns tavg_total_time. */
int lp_tavg_reset(lp_ctx *ctx);
int lp_tavg_compute(lp_ctx *ctx, double tavg_dt);
int lp_tavg_get(lp_ctx *ctx, int32_t q, void *out, size_t bytes, double *total_time);
/* io.f90:1643-1679 (spectrum), called on every level as avg_stats (io.f90:1846-1848,
 * E11 before its division by z) and post_spec (io.f90:1587-1594) do: out (host,
 * binary64) holds (nz-1) rows of nx/2, row k the y-average of the streamwise power of
 * level k, sum_y w_m |X_m|^2 / ny with X the forward transform of u(0:nx-1, y)/nx,
 * w_0 = 1/2 (real part), w_m = 1; no Nyquist term.  field: a volume of the context
 * (not LP_P, LP_USTAR); bin selects the size bin of a per-bin volume (LP_PCON ...),
 * 0 otherwise.  The transform runs in the field's own type. */
int lp_spectrum(lp_ctx *ctx, int32_t field, int32_t bin, double *out);
/* scalars_module_crossflow_mod.f90:1494-1786 (pollen_all_in_one2 with
 * pollen_RHS_calc2, calcbeta_pcon, step_pollen2) on the context state; needs npcon > 0. */
int lp_pollen_all_in_one2(lp_ctx *ctx);
/* The three parts of that step on their own (scalars_module_crossflow_mod.f90):
 * pollen_RHS_calc2 (:1894-4472; writes RHS_PCon, Rhs_pbe, magS(nz)), calcbeta_pcon
 * (:1795-1843; writes beta_pcon) and step_pollen2 (:4481-4508; advances PCon). */
int lp_pollen_rhs_calc2(lp_ctx *ctx);
int lp_calcbeta_pcon(lp_ctx *ctx);
int lp_step_pollen2(lp_ctx *ctx);
/* scalars_module_crossflow_mod.f90:4530-4597: beta(npcon,npcon) binary64, beta[i*n+j]
 * = probability that a drop of bin j breaks into bin i (host routine, no device). */
int lp_calculate_probability(const double *diameter, int32_t npcon, double sigma_s,
                             double d_min, double *beta);
/* One or more device-resident time steps of the hot loop (main.f90:207-1004).
 * Status, checked once per call after its last step (call it with nsteps = 1 to check
 * every step): LP_ERR_PIVOT if a pressure solve met a zero pivot, LP_ERR_NAN if a step
 * wrote a non-finite u, v or w, LP_ERR_CFL if a limit is set and the CFL of the final
 * state (timestep_conditions) exceeds it.  The steps of the call have all run. */
int lp_step(lp_ctx *ctx, int32_t nsteps);
/* The reference reports CFL and divergence every wbase steps (main.f90:902-926) and does
 * not stop on them; with cfl_max > 0 lp_step returns LP_ERR_CFL above it (0: off). */
int lp_set_cfl_limit(lp_ctx *ctx, double cfl_max);
/* Scan u, v, w, RHS, RHS_f, p, F_LM/MM/QN/NN, Cs_opt2 (and the droplet state) for
 * non-finite values: LP_ERR_NAN naming the first such field, else LP_OK. */
int lp_check_finite(lp_ctx *ctx);
/* The context's enum lp_precision.  The explicit-array routines take arrays of the
 * context's field type: binary64 for LP_FP64, binary32 for LP_MIXED. */
int lp_get_precision(lp_ctx *ctx, int32_t *precision);

/* ---------------- frozen-coefficient steps and the discrete adjoint ----------------
 * No counterpart in the reference (which has no adjoint).  With nut (a volume) and ust (a
 * plane), device arrays of the field type, lp_step runs frozen-coefficient steps: the
 * eddy viscosity Nu_t and the wall u* are the given ones instead of the SGS model's
 * (sgs_stag.f90:137-227) and the wall law's (wallstress.f90:145-162), so the step is a
 * quadratic map of (u, v, w, RHS_f).  Null for both restores the model. */
int lp_set_frozen(lp_ctx *ctx, const void *nut, const void *ust);
/* The transpose of that step's linearisation at the context's state (u, v, w before the
 * step, jt; jt == 1 without initu is the Euler start), binary64 only (LP_FP64 context;
 * the base states of an LP_MIXED run carry over, u, v, w being binary64 in both).  nut,
 * ust: the frozen coefficients of the step (binary64 device arrays).  ub (3 volumes
 * u, v, w) and rb (3 volumes RHSx..z), device arrays: on entry the adjoints of the step's
 * outputs, on return those of its inputs (ub of u, v, w, rb of RHS_f).  The context's
 * state is left unchanged; its first call allocates the adjoint's scratch. */
int lp_step_adjoint(lp_ctx *ctx, const double *nut, const double *ust, double *ub, double *rb);
/* lp_step_adjoint with droplets (required when npcon > 0; the base C is the context's
 * PCon): cb, rcb (npcon volumes each, binary64 device arrays), on entry the adjoints of
 * the step's PCon and RHS_PCon, on return those of the PCon and RHS_PCon it starts from
 * (pollen_all_in_one2, scalars_module_crossflow_mod.f90:1494-1786, with the transport
 * linear in C on the limiter branch the base C selects).  Inertia (PCon_acc) and the
 * breakup feedback are not differentiated: LP_ERR_ARG. */
int lp_step_adjoint2(lp_ctx *ctx, const double *nut, const double *ust, double *ub, double *rb, double *cb,
                     double *rcb);

/* ---------------- measurement ----------------
 * Kernel classes whose launches can be bracketed by HIP events on the context's
 * stream while lp_step runs (no counterpart in the reference; used by bench.py). */
enum lp_kernel_class {
  LP_K_ROW_R2C = 0,        /* row pass real -> half spectrum                        */
  LP_K_COL_OP = 1,         /* column FFT + spectral map + inverse column FFT          */
  LP_K_ROW_C2R = 2,        /* row pass half spectrum -> real                        */
  LP_K_LAG_RELAX = 3,      /* Germano contractions + Lagrangian relaxation           */
  LP_K_PRESS_TRIDIAG = 4,  /* per-mode binary64 Thomas solve                        */
  LP_K_POLLEN_RHS = 5,     /* droplet transport RHS of all bins (SMART + diffusion)   */
  LP_K_PBE = 6,            /* droplet breakup source of all bins                     */
  LP_K_CONVEC = 7,         /* 3/2-grid x back-transforms + cross products + x forward */
  LP_K_COL_PAD = 8,        /* column passes padding to / truncating from the 3/2 grid */
  LP_K_COL_DIV = 9,        /* column pass summing the x and y stress derivatives      */
  LP_K_NCLASSES = 10
};
/* Start (kernel_class >= 0) or stop (-1) event timing; starting clears the totals. */
int lp_prof_enable(lp_ctx *ctx, int32_t kernel_class);
/* Totals since lp_prof_enable: summed kernel time, launches, and the kernels'
 * compulsory bytes (each input read once, each output written once). */
int lp_prof_read(lp_ctx *ctx, double *total_ms, int64_t *launches, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
