/*
 * lesgo_oracle.h -- CPU binary64 restatement of the reference's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP implementation
 * and the timed CPU baseline; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in population_balance_amd/ links it.
 *
 * Layouts and configuration are those of include/lesgo_amd.h (all fields binary64).
 */
#ifndef LESGO_ORACLE_H
#define LESGO_ORACLE_H
#include "../include/lesgo_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(const lp_config *cfg);
void or_destroy(or_ctx *c);
double *or_field(or_ctx *c, int32_t field);
size_t or_field_elems(or_ctx *c, int32_t field);
void or_set_time(or_ctx *c, int32_t jt, int32_t jt_total);
void or_get_time(or_ctx *c, int32_t *jt, int32_t *jt_total);
int or_set_threads(int n);
const char *or_last_error(void);

/* 2-D plane transforms with FFTW2 rfftwnd in-place semantics (unnormalised). */
void or_rfft2_fwd(or_ctx *c, int big, const double *in, int pitch, double *out_cplx);
void or_rfft2_inv(or_ctx *c, int big, const double *in_cplx, double *out, int pitch);
/* 1-D complex transform of length n (sign -1 forward, +1 backward), unnormalised. */
void or_cfft1(int n, int sign, const double *in_cplx, double *out_cplx);

void or_init_wavenumber(or_ctx *c, double *kx, double *ky, double *k2);
void or_test_filter_init(or_ctx *c, double alpha, double *G);
void or_filt_da(or_ctx *c, double *f, double *dfdx, double *dfdy, int nlev);
void or_ddx(or_ctx *c, double *dfdx, const double *f, int nlev);
void or_ddy(or_ctx *c, double *dfdy, const double *f, int nlev);
void or_ddxy(or_ctx *c, double *dfdx, double *dfdy, const double *f, int nlev);
void or_ddz_uv(or_ctx *c, double *dfdz, const double *f);
void or_ddz_w(or_ctx *c, double *dfdz, const double *f);
void or_padd(or_ctx *c, double *u_big, const double *u, int nplanes);
void or_unpadd(or_ctx *c, double *cc, const double *cc_big, int nplanes);
void or_dealias1(or_ctx *c, const double *u, double *u_big, int nlev);
void or_dealias2(or_ctx *c, double *u, double *u_big, int nlev);
void or_test_filter(or_ctx *c, double *f, const double *G, int nplanes);
void or_divstress_uv(or_ctx *c, double *divt, const double *tx, const double *ty,
                     const double *tz);
void or_divstress_w(or_ctx *c, double *divt, const double *tx, const double *ty,
                    const double *tz);
int or_tridag_array(or_ctx *c, const double *a, const double *b, const double *cc,
                    const double *r, double *u);

/* routines on the context state */
int or_wallstress(or_ctx *c);
int or_sgs_stag(or_ctx *c);
int or_lagrange_sdep(or_ctx *c);
int or_interpolag_sdep(or_ctx *c);
int or_convec(or_ctx *c);
int or_std_dynamic(or_ctx *c);
void or_plane_dynamic_level(or_ctx *c, int model, int k, double *out);
void or_dyn_contract(or_ctx *c, int k, double *lm, double *mm);
int or_scaledep_dynamic(or_ctx *c);
int or_lagrange_ssim(or_ctx *c);
int or_interpolag_ssim(or_ctx *c);
int or_obukhov(or_ctx *c);
void or_obukhov_get(or_ctx *c, double *out);
int or_theta_all_in_one(or_ctx *c);
int or_press_stag_array(or_ctx *c);
int or_project(or_ctx *c);
void or_timestep_conditions(or_ctx *c, double *cfl, double *visc_stab);
void or_rmsdiv(or_ctx *c, double *rms);
void or_plane_stats(or_ctx *c, double *out);
double or_energy(or_ctx *c, double *ke);
void or_tavg_compute(or_ctx *c, double tavg_dt, double *acc);
int or_spectrum(or_ctx *c, int field, int bin, double *out);
int or_step(or_ctx *c, int nsteps);
int or_pollen_all_in_one2(or_ctx *c);
int or_pollen_rhs_calc2(or_ctx *c);
int or_calcbeta_pcon(or_ctx *c);
int or_step_pollen2(or_ctx *c);
void or_ke_dudt(or_ctx *c);
void or_calculate_probability(const double *d, int n, double sigma, double d_min, double *beta);
/* frozen-coefficient steps and their discrete adjoint */
void or_set_frozen(or_ctx *c, const double *nut, const double *ust);
int or_step_adjoint(or_ctx *c, const double *nut, const double *ust, double *ub, double *rb);
int or_step_adjoint2(or_ctx *c, const double *nut, const double *ust, double *ub, double *rb, double *cb,
                     double *rcb);

#ifdef __cplusplus
}
#endif
#endif
