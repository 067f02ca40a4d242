// lp_api.hip -- C ABI of include/lesgo_amd.h: context, device state, and the host-side
// orchestration of one LES time step (main.f90:207-826) on one HIP stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/lesgo_amd.h"
#include "lp_internal.hpp"
#include "lp_kernels.hpp"
#include "lp_droplets.hpp"
#include "lp_lag.hpp"
#include "lp_convec.hpp"
#include "lp_adjoint.hpp"
#include "lp_scalar.hpp"
#include "lp_sgs_alt.hpp"

using namespace lp;

static thread_local char g_err[512];
static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
namespace lp {
int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
}  // namespace lp
#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_err(LP_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                     __FILE__, __LINE__);                                                   \
  } while (0)
#define RCHK(expr)          \
  do {                      \
    int r_ = (expr);        \
    if (r_ != LP_OK) return r_; \
  } while (0)

// ------------------------------------------------------------------------------------
// twiddle registry (per device, per length)
// ------------------------------------------------------------------------------------
namespace {
struct TwEntry {
  void *d = nullptr, *f = nullptr;
};
std::mutex g_tw_mu;
std::map<std::pair<int, int>, TwEntry> g_tw;

int ensure_twiddle(int n) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, n);
  if (g_tw.count(key)) return LP_OK;
  std::vector<double> hd(2 * (size_t)n);
  std::vector<float> hf(2 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
    double c = (double)cosl(a), s = (double)-sinl(a);
    hd[2 * j] = c; hd[2 * j + 1] = s;
    hf[2 * j] = (float)cosl(a); hf[2 * j + 1] = (float)-sinl(a);
  }
  TwEntry e;
  HIPCHK(hipMalloc(&e.d, sizeof(double) * 2 * n));
  HIPCHK(hipMalloc(&e.f, sizeof(float) * 2 * n));
  HIPCHK(hipMemcpy(e.d, hd.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.f, hf.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
  HIPCHK(hipDeviceSynchronize());  // see dalloc: the tables are read on non-blocking streams
  g_tw[key] = e;
  return LP_OK;
}
}  // namespace

namespace lp {
const void *tw_ptr(int n, Dt dt) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw.find(std::make_pair(dev, n));
  if (it == g_tw.end()) return nullptr;
  return dt == F64 ? it->second.d : it->second.f;
}
}  // namespace lp

// ------------------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------------------
static const int kNumLag = 21;

struct lp_ctx {
  lp_config cfg;
  lp_grid g;
  int dev = 0;
  hipStream_t s = nullptr;
  // second stream of the step: the single-plane wall-stress passes beside the
  // z-derivatives and the pressure's wall/lid rows beside its H transforms (each between
  // ev_aux_fork and ev_aux), the convective term beside the SGS model (ev_fork after the
  // Lagrangian column pass .. ev_join before the RHS assembly)
  hipStream_t s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_aux_fork = nullptr, ev_aux = nullptr;
  bool prof_off = false;  // work on s2 that overlaps other kernels is not profiled
  void *cv_spec[2] = {};  // convec's own spectral scratch (divstress runs meanwhile)
  bool fork_pending = false;  // the step's convec is still to be started on s2
  Dt fdt = F64;  // field type
  void *f[LP_NFIELDS] = {};
  size_t bytes[LP_NFIELDS] = {};
  int esz[LP_NFIELDS] = {};
  double *kx = nullptr, *ky = nullptr, *k2 = nullptr, *G1 = nullptr, *G2 = nullptr;
  SpecTables tab{};
  // scratch
  void *spec[4] = {};         // small spectra (lh x ny x nz) of the field type
  void *specb[2] = {};        // big spectra (lh_big x ny2 x nz)
  // Internal spectra keep rows of sp (small) / sb (big) complex values: lh and lh_big
  // rounded up to 16, so that every row starts on a 128-byte boundary.
  long sp = 0, sb = 0;
  void *phat = nullptr;       // binary64 p-hat (0..nz) in the sp layout
  void *uvw_spec[3] = {};     // x-spectra of the filtered u, v, w left by the step's filt_da
  // binary32 context: the per-level values that u and v entered filt_da's transform
  // shifted by ([2][nz]); the convective pass adds them back to the kept spectra
  double *ushift = nullptr;
  double *nyq = nullptr;    // filt_da of binary64 velocities: A [nz][ny], B [nz][nx], C [nz]
  double *nyq_part = nullptr;  // their partial column sums [nz][kNyqSlices][nx]
  double *adj_drop = nullptr;  // droplet adjoint scratch: mu, ac, awi volumes + aui, avi faces
  double *tavg = nullptr;      // running time averages (tavg_nq binary64 volumes)
  int tavg_nq = 0;
  double tavg_time = 0.0;      // tavg_total_time
  double *d_pvol = nullptr;    // per bin: pi/6 d^3 (tavg's volume fractions)
  // potential temperature (cfg.scal.s_flag): obukhov's results (psi_m, psi_h, phi_m,
  // phi_h, u*, L, <sgs_t3(1)>, <theta(1)>), the per-level shift of theta's binary32
  // transform, calcbeta's level means and the six 3/2-grid fields of the advection
  double *d_obuk = nullptr, *tshift = nullptr, *d_tbar = nullptr;
  void *tbig[6] = {};
  // SGS models 2-4 (lp_sgs_alt.hpp): the 21 filter inputs, their 2-delta and (model 3)
  // 4-delta filtered volumes, model 3's plane sums [nz][10] and beta [nz]
  void *dyn_in = nullptr, *dyn_bar = nullptr, *dyn_hat = nullptr;
  double *dyn_sums = nullptr, *dyn_beta = nullptr;
  double lzo_ = 0.0;           // log(0.5 dz / zo_), zo_ = exp(mean(log z0)) (obukhov :1205)
  void *cvb = nullptr;        // convec: y-padded half spectra of u1 u2 u3 w1 w2 w3 [6][nz][ny2][cvp_pitch]
  void *cvp = nullptr;        // convec: half spectra of the 3/2-grid products [3][nz][ny2][cvp_pitch]
  long cv_pitch = 0;          // kx columns kept on the 3/2 grid (nx/2) rounded up to 16
  void *tmp[4] = {};          // field-type volumes
  void *S[6] = {};            // strain rates
  void *lspec = nullptr;      // x spectra of the 21 test-filter inputs [21][nz][ny][pad16(nc)]
  void *lbar = nullptr, *lhat = nullptr;  // their y-filtered spectra [21][nz][ny][pad16(ncut)]
  void *lcon[4] = {};         // Germano contractions L.M, M.M, Q.N, N.N (field-type volumes)
  double *Fi[4] = {};         // second buffer of the Lagrangian averages (exchanged with F)
  void *Hs[3] = {};           // pressure RHS spectra
  void *bw = nullptr, *tw = nullptr;  // wall / lid divergence spectra
  double *gam = nullptr;      // Thomas scratch (nz+1) x lh x ny
  // the pressure systems' elimination factors gam(j), 1/bet(j) ((nz+1) x lh x ny), fixed by
  // the grid: computed once in lp_create (k_press_factors)
  double *pgam = nullptr, *ibet = nullptr;
  void *pspec[3] = {};        // binary64 spectra for the p inverse
  // d_err[0]: zero pivot of a Thomas solve; d_err[1]: non-finite velocity written by a
  // step; d_err[2]: lp_check_finite's scratch flag
  int *d_err = nullptr;
  double cfl_max = 0.0;  // lp_set_cfl_limit (0: no check, as the reference only reports it)
  // frozen-coefficient steps (lp_set_frozen): Nu_t (volume) and the wall u* (plane) of a
  // base run, device arrays of the field type, used instead of the SGS model / wall law
  const void *frz_nut = nullptr, *frz_ust = nullptr;
  // lp_step_adjoint scratch (allocated on its first call): binary64 volumes, 3/2-grid
  // volumes and the transposed solve's lambda
  double *adj_block = nullptr;
  size_t adj_bytes = 0;
  double *d_stats = nullptr;  // lp_plane_stats: [nz-1][10]
  double *d_espec = nullptr;  // lp_spectrum: [nz-1][nx/2]
  double *d_part = nullptr;
  int nred = 0;
  int jt = 1, jt_total = 1;
  int F_LM_MM_init = 0, F_QN_NN_init = 0;
  // u_lag, v_lag, w_lag known to hold zeros (after their reset, or as allocated); cleared
  // by an accumulation, an upload, or handing out their device pointers
  bool lag_vel_zero = true;
  int ncut1 = 0, ncut2 = 0;  // leading kx columns where G_test / G_test_test are nonzero
  // droplet bins (npcon > 0)
  double *d_ws = nullptr, *d_qsrc = nullptr, *d_vfac = nullptr;  // per bin
  double *d_foot = nullptr;                    // normalised source footprint (nx*ny)
  PbeBin *d_bins = nullptr;
  double *d_pbeA = nullptr;  // [npc][npc] pair-transfer matrix of the breakup source
  double *ui0 = nullptr, *vi0 = nullptr;        // face velocities (ld, ny+1, nz)
  double *kcell = nullptr;                      // Kc_t on cell centres
  double kpref = 0.0, ustar3 = 0.0;
  // event timing of one kernel class (lp_prof_enable)
  int prof_class = -1;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  double prof_bytes = 0.0, prof_ms = 0.0;
  long prof_launches = 0;
};

// bracket a launch of kernel class `cls` with events when that class is profiled
// a free pair of timing events for class cls, when it is being profiled
static bool prof_reserve(lp_ctx *c, int cls) {
  if (c->prof_class != cls || c->prof_off) return false;
  if (c->ev_used + 2 > c->ev.size()) {
    for (int i = 0; i < 256; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return false;
      c->ev.push_back(e);
    }
  }
  return true;
}
static void prof_begin(lp_ctx *c, int cls) {
  if (prof_reserve(c, cls)) (void)hipEventRecord(c->ev[c->ev_used], c->s);
}
static void prof_end(lp_ctx *c, int cls, double bytes) {
  if (c->prof_class != cls || c->prof_off || c->ev_used + 2 > c->ev.size()) return;
  (void)hipEventRecord(c->ev[c->ev_used + 1], c->s);
  c->ev_used += 2;
  c->prof_bytes += bytes;
  c->prof_launches += 1;
}
// the pair for one launch that takes the events itself (hipExtLaunchKernel: no marker
// packets around it, so the measurement adds no idle time to the stream), then its account
static LaunchEv prof_attach(lp_ctx *c, int cls) {
  if (!prof_reserve(c, cls)) return LaunchEv{};
  return LaunchEv{c->ev[c->ev_used], c->ev[c->ev_used + 1]};
}
static void prof_attached(lp_ctx *c, const LaunchEv &ev, double bytes) {
  if (!ev.start) return;
  c->ev_used += 2;
  c->prof_bytes += bytes;
  c->prof_launches += 1;
}
// fold recorded event pairs into the running total (call after a stream sync)
static void prof_collect(lp_ctx *c) {
  for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]) == hipSuccess) c->prof_ms += ms;
  }
  c->ev_used = 0;
}

static size_t esz_of(Dt d) { return d == F64 ? 8 : 4; }
// a spectral row length rounded up so that rows start on 128-byte boundaries
static long pad16(long n) { return (n + 15) / 16 * 16; }
static long plane_el(const lp_ctx *c) { return (long)c->g.ld * c->g.ny; }
static long vol_el(const lp_ctx *c) { return plane_el(c) * c->g.nz; }
static Geo geo(const lp_ctx *c) {
  Geo g;
  g.nx = c->g.nx; g.ny = c->g.ny; g.nz = c->g.nz; g.ld = c->g.ld;
  g.np = plane_el(c);
  return g;
}
static dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

static bool field_is_f64(const lp_ctx *c, int id) {
  if (c->fdt == F64) return true;
  switch (id) {
    case LP_U: case LP_V: case LP_W:
    case LP_P: case LP_F_LM: case LP_F_MM: case LP_F_QN: case LP_F_NN:
    case LP_PCON: case LP_RHS_PCON: case LP_RHS_PCON_F: case LP_RHS_PBE:
    case LP_THETA: case LP_SGS_T3: case LP_PSI_M:
      return true;
    default:
      return false;
  }
}

// Every kernel and copy of a context runs on its own non-blocking streams, which are not
// ordered after work on the null stream: the zero fill (and any synchronous table upload,
// whose DMA may still be in flight when hipMemcpy returns from pageable memory) is waited
// for here, so that a buffer allocated while the context computes (the adjoint's scratch,
// the running averages) is not zeroed under the kernels that already write it.
static int dalloc(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) return set_err(LP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  e = hipMemset(*p, 0, bytes ? bytes : 16);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return set_err(LP_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
  return LP_OK;
}

extern "C" int lp_abi_version(void) { return LP_ABI_VERSION; }
extern "C" const char *lp_last_error(void) { return g_err; }

extern "C" int lp_grid_init(lp_grid *g, int32_t nx, int32_t ny, int32_t nz, double L_x, double L_y,
                            double L_z) {
  if (!g || nx < 4 || ny < 4 || nz < 3 || (nx & 1) || (ny & 1))
    return set_err(LP_ERR_ARG, "lp_grid_init: invalid sizes nx=%d ny=%d nz=%d", nx, ny, nz);
  std::memset(g, 0, sizeof *g);
  g->nx = nx; g->ny = ny; g->nz = nz;
  g->lh = nx / 2 + 1;
  g->ld = 2 * g->lh;
  g->nx2 = 3 * nx / 2;
  g->ny2 = 3 * ny / 2;
  g->lh_big = g->nx2 / 2 + 1;
  g->ld_big = 2 * g->lh_big;
  g->L_x = L_x; g->L_y = L_y; g->L_z = L_z;
  g->dx = L_x / nx;
  g->dy = L_y / ny;
  g->dz = L_z / (nz - 1);
  return LP_OK;
}

extern "C" void lp_config_default(lp_config *c) {
  std::memset(c, 0, sizeof *c);
  lp_grid_init(&c->grid, 32, 32, 33, 2.0 * M_PI, 2.0 * M_PI, 1.0);
  c->dt = 2e-4;
  c->precision = LP_FP64;
  c->model = 5;
  c->inilag = 1; c->initu = 0;
  c->cs_count = 1; c->dyn_init = 2;
  c->fp32_casts = 1;
  c->lbc_mom = LP_LBC_WALL;
  c->z0 = 1e-4; c->d0 = 0.0; c->vonk = 0.4;
  c->Co = 0.16; c->nnn = 2;
  c->use_mean_p_force = 0; c->mean_p_force = 1.0;
  c->cs2_init = 0.03; c->opftime = 1.5; c->filter_size = 1.0;
  // param.f90:96, 244-283 (theta off; Pr 0.4, SCAL_init 1, lbc 1, T_scale 300 K, g 9.81)
  lp_scalar_config &S = c->scal;
  S.s_flag = 0; S.passive = 0; S.scal_init = 1; S.initsc = 0; S.lbc = 1; S.ocean = 0; S.iboussinesq = 3;
  S.Pr = 0.4; S.wt_s = 0.0; S.T_scale = 300.0; S.dTdz_top = 0.0; S.g = 9.81; S.u_star = 1.0; S.z_i = 1.0;
}

static void host_wavenumbers(const lp_grid &g, std::vector<double> &kx, std::vector<double> &ky,
                             std::vector<double> &k2) {
  // fft.f90:51-76
  int lh = g.lh, ny = g.ny;
  kx.assign((size_t)lh * ny, 0.0); ky = kx; k2 = kx;
  for (int jy = 0; jy < ny; ++jy)
    for (int jx = 0; jx < lh; ++jx) {
      size_t i = (size_t)jy * lh + jx;
      double vx = (jx < lh - 1) ? (double)jx : 0.0;
      double vy = (double)(((jy + ny / 2) % ny) - ny / 2);
      if (jx == lh - 1 || jy == ny / 2) { vx = 0.0; vy = 0.0; }
      kx[i] = 2.0 * 3.1415926535897932384626433 / g.L_x * vx;
      ky[i] = 2.0 * 3.1415926535897932384626433 / g.L_y * vy;
      k2[i] = kx[i] * kx[i] + ky[i] * ky[i];
    }
}

static void host_filter(const lp_grid &g, const std::vector<double> &k2, double alpha, std::vector<double> &G) {
  // test_filtermodule.f90:29-57, ifilter = 1
  int lh = g.lh, ny = g.ny;
  G.assign((size_t)lh * ny, 0.0);
  double delta = alpha * std::sqrt(g.dx * g.dy);
  double pi = 3.1415926535897932384626433;
  double kc2 = (pi / delta) * (pi / delta);
  for (int jy = 0; jy < ny; ++jy)
    for (int jx = 0; jx < lh; ++jx) {
      size_t i = (size_t)jy * lh + jx;
      double v = 1.0 / ((double)g.nx * g.ny);
      if (k2[i] >= kc2) v = 0.0;
      if (jx == lh - 1 || jy == ny / 2) v = 0.0;
      G[i] = v;
    }
}

// scalars_module_crossflow_mod.f90:4530-4597: daughter-size probability from the
// surface-energy increase of each split (Tsouris-Tavlarides style).  beta[i*n+j] is the
// probability that a drop of bin j (> i) yields a drop of bin i.
static void host_breakup_probability(const double *d, int n, double sigma, double d_min, double *beta) {
  std::vector<double> cube(n), sq(n), e_max(n), e_min(n), e_split((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    cube[i] = std::pow(d[i], 3.0);
    sq[i] = std::pow(d[i], 2.0);
  }
  for (int i = 0; i < n; ++i) {
    e_max[i] = M_PI * sigma * 0.25992104 * sq[i];
    e_min[i] = M_PI * sigma * (std::pow(cube[i] - d_min * d_min * d_min, 2.0 / 3.0) + d_min * d_min - sq[i]);
    for (int j = i + 1; j < n; ++j)
      e_split[(size_t)i * n + j] = M_PI * sigma * (std::pow(cube[j] - cube[i], 2.0 / 3.0) + sq[i] - sq[j]);
  }
  for (int q = 0; q < n * n; ++q) beta[q] = 0.0;
  for (int j = 1; j < n; ++j) {
    double norm = 0.0;
    for (int m = 0; m < j; ++m) norm += e_min[j] + e_max[j] - e_split[(size_t)m * n + j];
    for (int i = 0; i < j; ++i) beta[(size_t)i * n + j] = (e_min[j] + e_max[j] - e_split[(size_t)i * n + j]) / norm;
  }
}

// :2167-2230 Ohnesorge-number fit coefficients of the breakage frequency of one bin
static PbeBin host_bin_fit(const lp_droplet_config &D, double dip) {
  PbeBin B;
  const double Oh = std::pow(D.rho_c * D.nu_d * D.nu_d / D.sigma_s / dip, 0.5);
  B.a = 2.392 * std::exp(-26.76 * Oh) + 2.877 * std::exp(-0.1244 * Oh);
  B.b = 0.45;
  B.c = -std::pow(10.0, (0.5446 * std::pow(Oh, -0.3776) / (1 + 12.67 * std::pow(Oh, 1.462))));
  B.d = -(4.172 * std::pow(Oh, -0.5492)) / (1 + 0.5079 * std::pow(Oh, -0.4879));
  // 0.7182 is a default-kind (binary32) literal in the reference
  const double l1 = std::log10(Oh + 1);
  B.e = std::pow(10.0, 0.4113 * std::exp(-55.94 * l1) + 0.5125 * std::exp(-(double)0.7182f * l1));
  B.dpow_tau = std::pow(dip, -0.6666667);
  B.dpow_re = std::pow(dip, 4.0 / 3.0);
  return B;
}

static int droplet_check(const lp_config &cf) {
  const lp_droplet_config &D = cf.drop;
  const lp_grid &g = cf.grid;
  if (D.npcon < 0 || D.npcon > 32) return set_err(LP_ERR_ARG, "npcon must be in 0..32 (got %d)", D.npcon);
  if (D.npcon == 0) return LP_OK;
  if (!D.diameter || !D.settling_vel || !D.Q_src)
    return set_err(LP_ERR_ARG, "droplets: diameter, settling_vel and Q_src must be given");
  for (int b = 0; b < D.npcon; ++b) {
    if (!(D.diameter[b] > 0.0)) return set_err(LP_ERR_ARG, "droplets: diameter[%d] must be > 0", b);
    if (b > 0 && !(D.diameter[b] > D.diameter[b - 1]))
      return set_err(LP_ERR_ARG, "droplets: diameters must increase with the bin index");
  }
  if (D.bc_x < 0 || D.bc_x > 2 || D.bc_y < 0 || D.bc_y > 2)
    return set_err(LP_ERR_ARG, "droplets: bc_x/bc_y must be 0 (periodic), 1 (inlet) or 2 (upwind ghosts)");
  if (!(D.Sc > 0.0) || !(D.u_star > 0.0) || !(D.z_i > 0.0) || !(D.g > 0.0) || !(D.nu_c > 0.0))
    return set_err(LP_ERR_ARG, "droplets: Sc, u_star, z_i, g and nu_c must be > 0");
  if (D.src_jt0 < D.src_jt1 && (D.src_k < 1 || D.src_k > g.nz - 2))
    return set_err(LP_ERR_ARG, "droplets: src_k must lie in 1..nz-2 (got %d)", D.src_k);
  if (D.src_jt0 < D.src_jt1 && !(D.src_sigma > 0.0)) return set_err(LP_ERR_ARG, "droplets: src_sigma must be > 0");
  return LP_OK;
}

static int upload_vec(void **dst, const void *src, size_t bytes) {
  RCHK(dalloc(dst, bytes));
  hipError_t e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipDeviceSynchronize();  // see dalloc
  if (e != hipSuccess) return set_err(LP_ERR_HIP, "table upload failed: %s", hipGetErrorString(e));
  return LP_OK;
}

// per-bin tables, source footprint and droplet scratch
static int droplet_setup(lp_ctx *c) {
  const lp_droplet_config &D = c->cfg.drop;
  const lp_grid &g = c->g;
  const int n = D.npcon;
  const long nv = vol_el(c);
  std::vector<double> vfac(n);
  const double gh = D.g * (D.z_i / (D.u_star * D.u_star));
  for (int b = 0; b < n; ++b) {  // :1795-1843 buoyancy weight of one bin (param.f90:665 volume)
    const double dd = D.diameter[b];
    const double V = M_PI / 6 * (dd * dd * dd);
    vfac[b] = gh * (D.rho_c / D.rho_w - 1.0) * V;
  }
  RCHK(upload_vec((void **)&c->d_ws, D.settling_vel, sizeof(double) * n));
  RCHK(upload_vec((void **)&c->d_qsrc, D.Q_src, sizeof(double) * n));
  RCHK(upload_vec((void **)&c->d_vfac, vfac.data(), sizeof(double) * n));
  std::vector<double> pvol(n);
  for (int b = 0; b < n; ++b) pvol[b] = M_PI / 6 * (D.diameter[b] * D.diameter[b] * D.diameter[b]);
  RCHK(upload_vec((void **)&c->d_pvol, pvol.data(), sizeof(double) * n));
  // :4222-4256 footprint normalised to unit sum
  std::vector<double> foot((size_t)g.nx * g.ny, 0.0);
  if (D.src_jt0 < D.src_jt1) {
    const double width = D.src_sigma * g.dx;
    const double area = 2.0 * M_PI * std::pow(width / g.dx, 2.0);
    double wsum = 0.0;
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double rx = ((i + 1) - D.src_i) * g.dx, ry = ((j + 1) - D.src_j) * g.dy;
        const double r2 = rx * rx + ry * ry;
        const double wv = std::exp(-std::pow(r2, D.src_ngau / 2) / (2.0 * std::pow(width, D.src_ngau)));
        foot[(size_t)j * g.nx + i] = wv;
        wsum += wv / area;
      }
    for (auto &x : foot) x = x / area / wsum;
  }
  RCHK(upload_vec((void **)&c->d_foot, foot.data(), sizeof(double) * foot.size()));
  // :2167-2230, :4356-4448 breakup tables
  std::vector<PbeBin> bins(n);
  for (int b = 0; b < n; ++b) bins[b] = host_bin_fit(D, D.diameter[b]);
  std::vector<double> beta((size_t)n * n);
  host_breakup_probability(D.diameter, n, D.sigma_s, D.d_min, beta.data());
  std::vector<PbePair> pairs;
  const double *d = D.diameter;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      PbePair q;
      q.i = i; q.j = j; q.lo = -1; q.hi = -1;
      q.beta = beta[(size_t)i * n + j];
      q.wlo = q.whi = 0.0;
      const double d0 = std::pow(d[j] * d[j] * d[j] - d[i] * d[i] * d[i], 1.0 / 3.0);
      if (d0 > d[0]) {  // split the complementary drop between the two nearest bins
        for (int cn = 1; cn < n; ++cn)
          if (!(d0 > d[cn])) {
            const double a1 = std::fabs(d[cn] - d0), b1 = std::fabs(d[cn - 1] - d0);
            const double ab = a1 + b1;
            q.lo = cn - 1;
            q.hi = cn;
            q.wlo = (a1 / ab) * ((d0 * d0 * d0) / (d[cn - 1] * d[cn - 1] * d[cn - 1]));
            q.whi = (b1 / ab) * ((d0 * d0 * d0) / (d[cn] * d[cn] * d[cn]));
            break;
          }
      } else {  // smaller than the first bin: its volume goes to bin 1
        q.lo = 0;
        q.wlo = (d0 * d0 * d0) / (d[0] * d[0] * d[0]);
      }
      pairs.push_back(q);
    }
  // the pair loop (:4356-4448) as a matrix: tr = gf(j) beta PCon(j) goes +1 into bin i,
  // -1 out of bin j and w_lo, w_hi into the complementary drop's bins
  std::vector<double> A((size_t)n * n, 0.0);
  for (const PbePair &q : pairs) {
    A[(size_t)q.i * n + q.j] += q.beta;
    A[(size_t)q.j * n + q.j] -= q.beta;
    if (q.lo >= 0) A[(size_t)q.lo * n + q.j] += q.beta * q.wlo;
    if (q.hi >= 0) A[(size_t)q.hi * n + q.j] += q.beta * q.whi;
  }
  RCHK(upload_vec((void **)&c->d_bins, bins.data(), sizeof(PbeBin) * n));
  RCHK(upload_vec((void **)&c->d_pbeA, A.data(), sizeof(double) * A.size()));
  c->kpref = D.k_b * std::pow(2.1, 0.5) * M_PI / 4.0 * 0.84;
  c->ustar3 = std::pow(D.u_star, 3.0);
  const size_t fb = sizeof(double) * (size_t)g.ld * (g.ny + 1) * g.nz;
  RCHK(dalloc((void **)&c->ui0, fb));
  RCHK(dalloc((void **)&c->vi0, fb));
  RCHK(dalloc((void **)&c->kcell, sizeof(double) * nv));
  return LP_OK;
}

extern "C" int lp_calculate_probability(const double *diameter, int32_t npcon, double sigma_s, double d_min,
                                        double *beta) {
  if (!diameter || !beta || npcon < 1) return set_err(LP_ERR_ARG, "lp_calculate_probability: bad argument");
  host_breakup_probability(diameter, npcon, sigma_s, d_min, beta);
  return LP_OK;
}

extern "C" int lp_destroy(lp_ctx *c) {
  if (!c) return LP_OK;
  (void)hipSetDevice(c->dev);
  if (c->s) (void)hipStreamSynchronize(c->s);
  auto fr = [](void *p) { if (p) (void)hipFree(p); };
  for (auto p : c->f) fr(p);
  fr(c->kx); fr(c->ky); fr(c->k2); fr(c->G1); fr(c->G2);
  for (auto p : c->spec) fr(p);
  for (auto p : c->specb) fr(p);
  fr(c->cvb); fr(c->cvp);
  for (auto p : c->uvw_spec) fr(p);
  for (auto p : c->tmp) fr(p);
  for (auto p : c->S) fr(p);
  fr(c->lspec); fr(c->lbar); fr(c->lhat);
  for (auto p : c->lcon) fr(p);
  for (auto p : c->Fi) fr(p);
  for (auto p : c->Hs) fr(p);
  fr(c->bw); fr(c->tw); fr(c->gam); fr(c->pgam); fr(c->ibet); fr(c->phat);
  for (auto p : c->pspec) fr(p);
  fr(c->d_err); fr(c->d_part); fr(c->d_stats); fr(c->d_espec); fr(c->ushift); fr(c->nyq); fr(c->nyq_part); fr(c->adj_block); fr(c->adj_drop); fr(c->tavg); fr(c->d_pvol);
  fr(c->d_obuk); fr(c->tshift); fr(c->d_tbar);
  for (auto p : c->tbig) fr(p);
  fr(c->dyn_in); fr(c->dyn_bar); fr(c->dyn_hat); fr(c->dyn_sums); fr(c->dyn_beta);
  fr(c->d_ws); fr(c->d_qsrc); fr(c->d_vfac); fr(c->d_foot); fr(c->d_bins); fr(c->d_pbeA);
  fr(c->ui0); fr(c->vi0); fr(c->kcell);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  for (auto p : c->cv_spec) fr(p);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_aux_fork) (void)hipEventDestroy(c->ev_aux_fork);
  if (c->ev_aux) (void)hipEventDestroy(c->ev_aux);
  if (c->s2 && c->s2 != c->s) (void)hipStreamDestroy(c->s2);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return LP_OK;
}

extern "C" int lp_create(const lp_config *cfg, lp_ctx **out) {
  if (!cfg || !out) return set_err(LP_ERR_ARG, "lp_create: null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(LP_ERR_NODEVICE, "lp_create: no HIP device visible");
  const lp_grid &g = cfg->grid;
  if (!fft_len_supported(g.nx) || !fft_len_supported(g.ny) || !fft_len_supported(g.nx2) ||
      !fft_len_supported(g.ny2))
    return set_err(LP_ERR_ARG, "lp_create: unsupported plane size %dx%d (supported: 32,64,96,128,192,256)",
                   g.nx, g.ny);
  if (g.ld != 2 * (g.nx / 2 + 1) || g.nz < 4) return set_err(LP_ERR_ARG, "lp_create: grid not initialised");
  if (cfg->model < 1 || cfg->model > 5)
    return set_err(LP_ERR_ARG, "lp_create: SGS model %d not implemented (1 to 5)", cfg->model);
  if (cfg->lbc_mom != LP_LBC_WALL && cfg->lbc_mom != LP_LBC_STRESS_FREE)
    return set_err(LP_ERR_ARG, "invalid lbc_mom");
  if (cfg->cs_count < 1) return set_err(LP_ERR_ARG, "cs_count must be >= 1");
  RCHK(droplet_check(*cfg));
  if (cfg->scal.s_flag) {
    const lp_scalar_config &S = cfg->scal;
    if (S.lbc != 1)
      return set_err(LP_ERR_ARG, "lp_create: theta lbc %d not implemented (1: prescribed surface flux; lbc 0 reads "
                                 "z_os, which the reference never sets)", S.lbc);
    if (!(S.Pr > 0) || !(S.T_scale > 0) || !(S.u_star > 0) || !(S.z_i > 0))
      return set_err(LP_ERR_ARG, "lp_create: theta needs Pr, T_scale, u_star, z_i > 0");
    if (S.ocean && (S.iboussinesq < 1 || S.iboussinesq > 3))
      return set_err(LP_ERR_ARG, "lp_create: invalid iboussinesq %d", S.iboussinesq);
  }
  lp_ctx *c = new lp_ctx();
  c->cfg = *cfg;
  c->g = g;
  c->fdt = cfg->precision == LP_MIXED ? F32 : F64;
  int rc = LP_OK;
#define CK(x)                      \
  do {                             \
    rc = (x);                      \
    if (rc != LP_OK) {             \
      lp_destroy(c);               \
      return rc;                   \
    }                              \
  } while (0)
  if (hipGetDevice(&c->dev) != hipSuccess) { delete c; return set_err(LP_ERR_HIP, "hipGetDevice failed"); }
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(LP_ERR_HIP, "hipStreamCreate failed");
  }
  // LP_SERIAL_STREAMS=1 (profiling): the second stream's work is issued on the first, so
  // that every kernel of a step runs alone and its trace duration is its own
  const char *serial = std::getenv("LP_SERIAL_STREAMS");
  const bool one_stream = serial && serial[0] == '1';
  if (one_stream) c->s2 = c->s;
  if ((!one_stream && hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess) ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_aux_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_aux, hipEventDisableTiming) != hipSuccess) {
    lp_destroy(c);
    return set_err(LP_ERR_HIP, "stream/event creation failed");
  }
  int lens[] = {g.nx / 2, g.nx, g.ny, g.nx2 / 2, g.nx2, g.ny2};
  for (int n : lens) CK(ensure_twiddle(n));
  long np = plane_el(c), nv = vol_el(c);
  for (int i = 0; i < LP_NFIELDS; ++i) {
    long n = nv;
    if (i == LP_P) n = np * (g.nz + 1);
    if (i == LP_USTAR || i == LP_PSI_M) n = np;
    if (i >= LP_THETA && i <= LP_BETA_SCAL && !cfg->scal.s_flag) n = 1;
    if (i == LP_PCON || i == LP_RHS_PCON || i == LP_RHS_PCON_F || i == LP_RHS_PBE)
      n = cfg->drop.npcon > 0 ? nv * cfg->drop.npcon : 1;
    if (i == LP_BETA_PCON && cfg->drop.npcon == 0) n = 1;
    c->esz[i] = field_is_f64(c, i) ? 8 : 4;
    c->bytes[i] = (size_t)n * c->esz[i];
    CK(dalloc(&c->f[i], c->bytes[i]));
  }
  std::vector<double> kx, ky, k2, G1, G2;
  host_wavenumbers(g, kx, ky, k2);
  host_filter(g, k2, 2.0 * cfg->filter_size, G1);
  host_filter(g, k2, 4.0 * cfg->filter_size, G2);
  size_t tb = sizeof(double) * kx.size();
  CK(dalloc((void **)&c->kx, tb)); CK(dalloc((void **)&c->ky, tb)); CK(dalloc((void **)&c->k2, tb));
  CK(dalloc((void **)&c->G1, tb)); CK(dalloc((void **)&c->G2, tb));
  if (hipMemcpy(c->kx, kx.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->ky, ky.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->k2, k2.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->G1, G1.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->G2, G2.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {  // see dalloc
    lp_destroy(c);
    return set_err(LP_ERR_HIP, "table upload failed");
  }
  c->tab.kx = c->kx; c->tab.ky = c->ky; c->tab.G1 = c->G1; c->tab.G2 = c->G2;
  auto cut = [&](const std::vector<double> &G) {
    int n = 1;
    for (int jy = 0; jy < g.ny; ++jy)
      for (int jx = 0; jx < g.lh; ++jx)
        if (G[(size_t)jy * g.lh + jx] != 0.0 && jx + 1 > n) n = jx + 1;
    return n;
  };
  c->ncut1 = cut(G1);
  c->ncut2 = cut(G2);
  c->tab.lh = g.lh; c->tab.ny = g.ny;
  size_t fe = esz_of(c->fdt);
  c->sp = (g.lh + 15) / 16 * 16;
  c->sb = (g.lh_big + 15) / 16 * 16;
  size_t spec_b = (size_t)c->sp * g.ny * g.nz * 2 * fe;
  size_t specb_b = (size_t)c->sb * g.ny2 * g.nz * 2 * fe;
  size_t vol_b = (size_t)nv * fe;
  for (auto &p : c->spec) CK(dalloc(&p, spec_b));
  for (auto &p : c->specb) CK(dalloc(&p, specb_b));
  c->cv_pitch = pad16(g.nx / 2);
  const size_t cv_field = (size_t)c->cv_pitch * g.ny2 * g.nz * 2 * fe;
  CK(dalloc(&c->cvb, 6 * cv_field));
  for (auto &p : c->uvw_spec) CK(dalloc(&p, spec_b));
  for (auto &p : c->cv_spec) CK(dalloc(&p, spec_b));
  CK(dalloc(&c->cvp, 3 * cv_field));
  for (auto &p : c->tmp) CK(dalloc(&p, vol_b));
  for (auto &p : c->S) CK(dalloc(&p, vol_b));
  if (cfg->model == 5) {
    const size_t pl = (size_t)kNumLag * g.ny * g.nz * 2 * fe;  // one kx column of all 21 fields
    const int nc = c->ncut1 > c->ncut2 ? c->ncut1 : c->ncut2;
    CK(dalloc(&c->lspec, pl * pad16(nc)));
    CK(dalloc(&c->lbar, pl * pad16(c->ncut1)));
    CK(dalloc(&c->lhat, pl * pad16(c->ncut2)));
    for (auto &p : c->lcon) CK(dalloc(&p, fe * nv));
    for (auto &p : c->Fi) CK(dalloc((void **)&p, sizeof(double) * nv));
  } else if (cfg->model >= 2) {
    CK(dalloc(&c->dyn_in, (size_t)kDynIn * vol_b));
    CK(dalloc(&c->dyn_bar, (size_t)kDynIn * vol_b));
    if (cfg->model == 3) {
      CK(dalloc(&c->dyn_hat, (size_t)kDynIn * vol_b));
      CK(dalloc((void **)&c->dyn_sums, sizeof(double) * 10 * (size_t)g.nz));
      CK(dalloc((void **)&c->dyn_beta, sizeof(double) * (size_t)g.nz));
    }
    if (cfg->model == 4)
      for (int q = 0; q < 2; ++q) CK(dalloc((void **)&c->Fi[q], sizeof(double) * nv));
  }
  for (auto &p : c->Hs) CK(dalloc(&p, spec_b));
  CK(dalloc(&c->bw, (size_t)c->sp * g.ny * 2 * fe));
  CK(dalloc(&c->tw, (size_t)c->sp * g.ny * 2 * fe));
  CK(dalloc((void **)&c->gam, sizeof(double) * (size_t)g.lh * g.ny * (g.nz + 1)));
  CK(dalloc((void **)&c->pgam, sizeof(double) * (size_t)g.lh * g.ny * (g.nz + 1)));
  CK(dalloc((void **)&c->ibet, sizeof(double) * (size_t)g.lh * g.ny * (g.nz + 1)));
  {
    const long nmodes = (long)g.lh * g.ny;
    hipLaunchKernelGGL(k_press_factors, grid1(nmodes), dim3(256), 0, c->s, (const double *)c->kx,
                       (const double *)c->ky, c->pgam, c->ibet, g.lh, g.ny, g.nz, g.dz);
    CK(hipGetLastError() == hipSuccess && hipStreamSynchronize(c->s) == hipSuccess
           ? LP_OK
           : set_err(LP_ERR_HIP, "lp_create: pressure factors"));
  }
  for (auto &p : c->pspec) CK(dalloc(&p, sizeof(double) * 2 * (size_t)c->sp * g.ny * (g.nz + 1)));
  CK(dalloc(&c->phat, sizeof(double) * 2 * (size_t)c->sp * g.ny * (g.nz + 1)));
  CK(dalloc((void **)&c->d_err, sizeof(int) * 4));
  CK(dalloc((void **)&c->ushift, sizeof(double) * 2 * (size_t)g.nz));
  CK(dalloc((void **)&c->nyq, sizeof(double) * (size_t)g.nz * (g.ny + g.nx + 1)));
  CK(dalloc((void **)&c->nyq_part, sizeof(double) * (size_t)g.nz * kNyqSlices * g.nx));
  CK(dalloc((void **)&c->d_stats, sizeof(double) * 10 * (size_t)(g.nz - 1)));
  if (cfg->scal.s_flag) {
    CK(dalloc((void **)&c->d_obuk, sizeof(double) * 8));
    CK(dalloc((void **)&c->tshift, sizeof(double) * (size_t)g.nz));
    CK(dalloc((void **)&c->d_tbar, sizeof(double) * (size_t)g.nz));
    for (auto &p : c->tbig) CK(dalloc(&p, (size_t)g.ld_big * g.ny2 * g.nz * fe));
    const double ob0[8] = {0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0};
    CK(hipMemcpy(c->d_obuk, ob0, sizeof ob0, hipMemcpyHostToDevice) == hipSuccess
           ? LP_OK
           : set_err(LP_ERR_HIP, "lp_create: obukhov initial values"));
    double slz = 0.0;  // :1205 sum(dlog(zo(1:nx,1:ny))) in order, zo uniform
    for (long q = 0; q < (long)g.nx * g.ny; ++q) slz += std::log(cfg->z0);
    c->lzo_ = std::log(0.5 * g.dz / std::exp(slz / ((double)g.nx * g.ny)));
  }
  CK(dalloc((void **)&c->d_espec, sizeof(double) * (size_t)(g.nx / 2) * (g.nz - 1)));
  c->nred = 512;
  CK(dalloc((void **)&c->d_part, sizeof(double) * 3 * c->nred));
  if (cfg->drop.npcon > 0) {
    rc = droplet_setup(c);  // reads the caller's bin arrays, which are not kept
    c->cfg.drop.diameter = c->cfg.drop.settling_vel = c->cfg.drop.Q_src = nullptr;
    if (rc != LP_OK) {
      lp_destroy(c);
      return rc;
    }
  }
#undef CK
  *out = c;
  return LP_OK;
}

extern "C" int lp_sync(lp_ctx *c) {
  if (!c) return set_err(LP_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  // s2 work is always joined into s, except after an error inside a step
  HIPCHK(hipStreamSynchronize(c->s2));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}

extern "C" int lp_field_info(lp_ctx *c, int32_t field, void **dev_ptr, size_t *bytes, int32_t *elem_size) {
  if (!c || field < 0 || field >= LP_NFIELDS) return set_err(LP_ERR_ARG, "lp_field_info: bad field %d", field);
  if (dev_ptr && (field == LP_U_LAG || field == LP_V_LAG || field == LP_W_LAG)) c->lag_vel_zero = false;
  if (dev_ptr) *dev_ptr = c->f[field];
  if (bytes) *bytes = c->bytes[field];
  if (elem_size) *elem_size = c->esz[field];
  return LP_OK;
}

extern "C" int lp_upload(lp_ctx *c, int32_t field, const void *host, size_t bytes) {
  if (!c || field < 0 || field >= LP_NFIELDS || !host) return set_err(LP_ERR_ARG, "lp_upload: bad argument");
  if (bytes != c->bytes[field])
    return set_err(LP_ERR_ARG, "lp_upload: field %d expects %zu bytes, got %zu", field, c->bytes[field], bytes);
  HIPCHK(hipSetDevice(c->dev));
  if (field == LP_U_LAG || field == LP_V_LAG || field == LP_W_LAG) c->lag_vel_zero = false;
  HIPCHK(hipMemcpyAsync(c->f[field], host, bytes, hipMemcpyDefault, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}

extern "C" int lp_download(lp_ctx *c, int32_t field, void *host, size_t bytes) {
  if (!c || field < 0 || field >= LP_NFIELDS || !host) return set_err(LP_ERR_ARG, "lp_download: bad argument");
  if (bytes != c->bytes[field])
    return set_err(LP_ERR_ARG, "lp_download: field %d holds %zu bytes, asked %zu", field, c->bytes[field], bytes);
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipMemcpyAsync(host, c->f[field], bytes, hipMemcpyDefault, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}

extern "C" int lp_set_time(lp_ctx *c, int32_t jt, int32_t jt_total) {
  if (!c) return set_err(LP_ERR_ARG, "null context");
  c->jt = jt;
  c->jt_total = jt_total;
  return LP_OK;
}
extern "C" int lp_get_time(lp_ctx *c, int32_t *jt, int32_t *jt_total) {
  if (!c) return set_err(LP_ERR_ARG, "null context");
  if (jt) *jt = c->jt;
  if (jt_total) *jt_total = c->jt_total;
  return LP_OK;
}

// ------------------------------------------------------------------------------------
// spectral pipelines (all asynchronous on c->s)
// ------------------------------------------------------------------------------------
// compulsory bytes of the FFT passes (input read once, output written once)
static double bytes_r2c(int nx, int rows, int nlev, int ncols_out, size_t e) {
  return (double)rows * nlev * ((double)nx * e + (double)ncols_out * 2 * e);
}
static double bytes_c2r(int nx, int rows, int nlev, int ncols_in, size_t ei, size_t eo) {
  return (double)rows * nlev * ((double)ncols_in * 2 * ei + (double)nx * eo);
}
static double bytes_col(int nyi, int nyo, int ncols, int nlev, int nout, size_t e) {
  return (double)ncols * nlev * 2 * e * ((double)nyi + (double)nout * nyo);
}

// ncols_out < lh writes only the leading kx columns (a spectrum that the caller knows to
// be zero beyond them)
// ri: the rows' type and shift (RowIn; default: rows of the field type, no shift)
static int r2c_small(lp_ctx *c, const void *in, int nlev, double scale, void *spec, int ncols_out = 0,
                     RowIn ri = {}) {
  const lp_grid &g = c->g;
  if (ncols_out <= 0) ncols_out = g.lh;
  const size_t ei = ri.din < 0 ? esz_of(c->fdt) : esz_of((Dt)ri.din);
  prof_begin(c, LP_K_ROW_R2C);
  HIPCHK(row_r2c(c->fdt, g.nx / 2, in, g.ld, plane_el(c), g.ny, nlev, scale, spec, c->sp,
                 c->sp * g.ny, ncols_out, c->s, ri));
  prof_end(c, LP_K_ROW_R2C, (double)g.ny * nlev * ((double)g.nx * ei + (double)ncols_out * 2 * esz_of(c->fdt)));
  return LP_OK;
}
// ikx: c->kx (its first row is kx of every column) to store the x-derivative instead;
// shift: per-plane values added back in binary64 (dto F64 from a binary32 spectrum)
static int c2r_small(lp_ctx *c, const void *spec, int ncols, int nlev, void *out, Dt dto,
                     const double *ikx = nullptr, const double *shift = nullptr) {
  const lp_grid &g = c->g;
  prof_begin(c, LP_K_ROW_C2R);
  HIPCHK(row_c2r(c->fdt, dto, g.nx / 2, spec, c->sp, c->sp * g.ny, ncols, g.ny, nlev, out, g.ld,
                 plane_el(c), c->s, ikx, shift));
  prof_end(c, LP_K_ROW_C2R, bytes_c2r(g.nx, g.ny, nlev, ncols, esz_of(c->fdt), esz_of(dto)));
  return LP_OK;
}
// column map on small spectra: in -> nout outputs (same grid)
static int col_small(lp_ctx *c, const void *in, int nlev, int fwd, int inv, int nout, ColOut *o, int ncols = 0) {
  const lp_grid &g = c->g;
  const long ps = c->sp * g.ny;
  if (ncols <= 0) ncols = g.lh;
  const LaunchEv ev = prof_attach(c, LP_K_COL_OP);
  HIPCHK(col_op(c->fdt, g.ny, g.ny, 0, fwd, inv, in, c->sp, ps, ncols, nlev, nout, o, c->sp, ps, c->tab, c->s,
                ev));
  prof_attached(c, ev, bytes_col(g.ny, g.ny, ncols, nlev, nout, esz_of(c->fdt)));
  return LP_OK;
}

// one input field -> derivative / filter outputs (multipliers), each output real field
static int spectral_apply(lp_ctx *c, const void *f, int nlev, double scale, int nout, const int *mults,
                          void *const *outs, RowIn ri = {}) {
  RCHK(r2c_small(c, f, nlev, scale, c->spec[0], 0, ri));
  ColOut o[3];
  for (int j = 0; j < nout; ++j) o[j] = ColOut{c->spec[j], mults[j]};
  RCHK(col_small(c, c->spec[0], nlev, 1, 1, nout, o));
  for (int j = 0; j < nout; ++j) RCHK(c2r_small(c, c->spec[j], c->g.lh, nlev, outs[j], c->fdt));
  return LP_OK;
}

// dealias1.f90: small field -> 3/2-grid physical field
static int to_big(lp_ctx *c, const void *f, int nlev, void *out_big, RowIn ri = {}) {
  const lp_grid &g = c->g;
  double cst = 1.0 / ((double)g.nx * g.ny);
  RCHK(r2c_small(c, f, nlev, cst, c->spec[0], 0, ri));
  ColOut o{c->specb[0], M_ONE};
  const long ps = c->sp * g.ny, pb = c->sb * g.ny2;
  HIPCHK(col_op(c->fdt, g.ny, g.ny2, 1, 1, 1, c->spec[0], c->sp, ps, g.nx / 2, nlev, 1, &o, c->sb, pb,
                c->tab, c->s));
  HIPCHK(row_c2r(c->fdt, c->fdt, g.nx2 / 2, c->specb[0], c->sb, pb, g.nx / 2, g.ny2, nlev, out_big,
                 g.ld_big, (long)g.ld_big * g.ny2, c->s));
  return LP_OK;
}

// dealias2.f90 core / convec.f90:161-170: 3/2-grid field (already scaled) -> small field
static int from_big(lp_ctx *c, const void *fb, int nlev, double scale, void *out) {
  const lp_grid &g = c->g;
  const long pb = c->sb * g.ny2, ps = c->sp * g.ny;
  HIPCHK(row_r2c(c->fdt, g.nx2 / 2, fb, g.ld_big, (long)g.ld_big * g.ny2, g.ny2, nlev, scale, c->specb[1],
                 c->sb, pb, g.nx / 2, c->s));
  ColOut o{c->spec[1], M_ONE};
  HIPCHK(col_op(c->fdt, g.ny2, g.ny, 2, 1, 1, c->specb[1], c->sb, pb, g.nx / 2, nlev, 1, &o, c->sp, ps,
                c->tab, c->s));
  RCHK(c2r_small(c, c->spec[1], g.nx / 2, nlev, out, c->fdt));
  return LP_OK;
}

#define FDISPATCH(c, T, ...)      \
  do {                            \
    if ((c)->fdt == F64) {        \
      using T = double;           \
      __VA_ARGS__;                \
    } else {                      \
      using T = float;            \
      __VA_ARGS__;                \
    }                             \
  } while (0)

// ------------------------------------------------------------------------------------
// routines on the context state
// ------------------------------------------------------------------------------------
// keep (optional): receives the x-spectrum of the filtered f (kx < lh), which the
// convection's 3/2 padding of the same field starts from.  kx is uniform along a column,
// so the (kx, y) spectrum of df/dx is i kx times that of f: the column pass produces f
// and df/dy only, and df/dx is the back-transform of f's spectrum times i kx.
// vel: f is a binary64 velocity of the context (in a binary32 context the transforms
// still run in binary32); shift (binary32 context only, with vel): the planes enter the
// transform less their first value, recorded in shift[plane] and added back to the
// filtered f in binary64 (filt_da(f - s) + s = filt_da(f), since a constant has only the
// kx = ky = 0 mode).  keep then holds the spectrum of f - s.
static int do_filt_da(lp_ctx *c, void *f, void *dfdx, void *dfdy, int nlev, void *keep = nullptr, bool vel = false,
                      double *shift = nullptr) {
  double cst = 1.0 / ((double)c->g.nx * c->g.ny);
  RowIn ri;
  if (vel) ri.din = F64;
  if (vel && shift && c->fdt == F32) {
    ri.shift = true;
    ri.shift_out = shift;
  } else {
    shift = nullptr;
  }
  RCHK(r2c_small(c, f, nlev, cst, c->spec[0], 0, ri));
  void *s0 = keep ? keep : c->spec[0];
  ColOut o[2] = {{s0, M_ONE}, {c->spec[2], M_IKY}};
  RCHK(col_small(c, c->spec[0], nlev, 1, 1, 2, o));
  if (vel && c->fdt == F32) {
    // the filtered velocity in binary64: its oddball modes subtracted exactly (a binary32
    // round trip would put about 3e-7 of the fluctuations' size back into u, v, w at
    // every step)
    Geo g = geo(c);
    double *A = c->nyq, *B = A + (size_t)c->g.nz * g.ny, *Cc = B + (size_t)c->g.nz * g.nx;
    if (g.nx > kNyqMax || g.ny > kNyqMax) return set_err(LP_ERR_ARG, "filt_da: nx, ny above %d", kNyqMax);
    hipLaunchKernelGGL(k_nyq_part, dim3((unsigned)(nlev * kNyqSlices)), dim3(256), 0, c->s, (const double *)f, A,
                       c->nyq_part, g);
    hipLaunchKernelGGL(k_nyq_fin, dim3((unsigned)nlev), dim3(256), 0, c->s, (const double *)A,
                       (const double *)c->nyq_part, B, Cc, g);
    hipLaunchKernelGGL(k_nyq_remove, grid1(g.np * nlev), dim3(256), 0, c->s, (double *)f, (const double *)A,
                       (const double *)B, (const double *)Cc, g, nlev);
    HIPCHK(hipGetLastError());
  } else {
    RCHK(c2r_small(c, s0, c->g.lh, nlev, f, vel ? F64 : c->fdt, nullptr, shift));
  }
  RCHK(c2r_small(c, s0, c->g.lh, nlev, dfdx, c->fdt, c->kx));
  return c2r_small(c, c->spec[2], c->g.lh, nlev, dfdy, c->fdt);
}

// vel: f is a binary64 velocity of the context (the difference is formed in binary64)
static int do_ddz_uv(lp_ctx *c, void *dfdz, const void *f, bool vel = false) {
  Geo g = geo(c);
  if (vel)
    FDISPATCH(c, T, hipLaunchKernelGGL((k_ddz_uv<T, double>), grid1(g.np * g.nz), dim3(256), 0, c->s, (T *)dfdz,
                                       (const double *)f, g, 1.0 / c->g.dz));
  else
    FDISPATCH(c, T, hipLaunchKernelGGL(k_ddz_uv<T>, grid1(g.np * g.nz), dim3(256), 0, c->s, (T *)dfdz,
                                       (const T *)f, g, (T)(1.0 / c->g.dz)));
  HIPCHK(hipGetLastError());
  return LP_OK;
}
static int do_ddz_w(lp_ctx *c, void *dfdz, const void *f, bool vel = false) {
  Geo g = geo(c);
  if (vel)
    FDISPATCH(c, T, hipLaunchKernelGGL((k_ddz_w<T, double>), grid1(g.np * g.nz), dim3(256), 0, c->s, (T *)dfdz,
                                       (const double *)f, g, 1.0 / c->g.dz));
  else
    FDISPATCH(c, T, hipLaunchKernelGGL(k_ddz_w<T>, grid1(g.np * g.nz), dim3(256), 0, c->s, (T *)dfdz,
                                       (const T *)f, g, (T)(1.0 / c->g.dz)));
  HIPCHK(hipGetLastError());
  return LP_OK;
}

static int do_wallstress(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  Geo g = geo(c);
  if (cf.lbc_mom == LP_LBC_STRESS_FREE) {  // wallstress.f90:200-205
    // fill kernels rather than hipMemsetAsync: this may run on the second stream, and the
    // runtime's memset there was seen to serialise the two streams
    const int zf[4] = {LP_TXZ, LP_TYZ, LP_DUDZ, LP_DVDZ};
    for (int q = 0; q < 4; ++q)
      FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(g.np), dim3(256), 0, c->s, (T *)c->f[zf[q]], g.np, T(0)));
    HIPCHK(hipGetLastError());
    return LP_OK;
  }
  // wallstress.f90:103-104,139-140: test-filtered copies of u(1), v(1)
  const int m[1] = {M_G1};
  void *o1[1] = {c->tmp[0]}, *o2[1] = {c->tmp[1]};
  RowIn vel;
  vel.din = F64;
  RCHK(spectral_apply(c, c->f[LP_U], 1, 1.0, 1, m, o1, vel));
  RCHK(spectral_apply(c, c->f[LP_V], 1, 1.0, 1, m, o2, vel));
  double zl = 0.5 * c->g.dz - cf.d0;
  double denom = std::log(zl / cf.z0);  // less psi_m (a plane from obukhov with theta on)
  if (c->frz_ust) {  // frozen coefficient: the base run's u*
    FDISPATCH(c, T, hipLaunchKernelGGL(k_wallstress_frozen<T>, grid1(g.np), dim3(256), 0, c->s,
                                       (const T *)c->tmp[0], (const T *)c->tmp[1], (const double *)c->f[LP_U],
                                       (const double *)c->f[LP_V], (T *)c->f[LP_TXZ], (T *)c->f[LP_TYZ],
                                       (T *)c->f[LP_DUDZ], (T *)c->f[LP_DVDZ], (T *)c->f[LP_USTAR],
                                       (const T *)c->frz_ust, g, denom, zl, cf.vonk));
    HIPCHK(hipGetLastError());
    return LP_OK;
  }
  FDISPATCH(c, T, hipLaunchKernelGGL(k_wallstress<T>, grid1(g.np), dim3(256), 0, c->s, (const T *)c->tmp[0],
                                     (const T *)c->tmp[1], (const double *)c->f[LP_U], (const double *)c->f[LP_V],
                                     (T *)c->f[LP_TXZ], (T *)c->f[LP_TYZ], (T *)c->f[LP_DUDZ],
                                     (T *)c->f[LP_DVDZ], (T *)c->f[LP_USTAR], g, denom, zl, cf.vonk,
                                     (const double *)(cf.scal.s_flag ? c->f[LP_PSI_M] : nullptr),
                                     (const double *)c->d_obuk));
  HIPCHK(hipGetLastError());
  return LP_OK;
}

// ------------------------------------------------------------------------------------
// potential temperature (scalars_module_crossflow_mod.f90, S_FLAG)
// ------------------------------------------------------------------------------------
// :1170-1403 obukhov (homogeneous surface): one workgroup over level 1
static int do_obukhov(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  const lp_scalar_config &S = cf.scal;
  Geo g = geo(c);
  ObkP P;
  P.stab = c->jt >= S.scal_init ? 1 : 0;
  P.zero_psi = (!P.stab && !S.initsc) ? 1 : 0;
  P.set_t3 = (P.stab && c->jt == S.scal_init && !S.initsc) ? 1 : 0;
  P.t3val = S.wt_s / S.u_star / S.T_scale;
  P.vonk = cf.vonk;
  P.dz = c->g.dz;
  P.lz0 = std::log(0.5 * c->g.dz / cf.z0);
  P.lzo_ = c->lzo_;
  P.g_ = S.g / (S.u_star * S.u_star / S.z_i);
  FDISPATCH(c, T, hipLaunchKernelGGL(k_obukhov<T>, dim3(1), dim3(1024), 0, c->s, (const double *)c->f[LP_U],
                                     (const double *)c->f[LP_V], (const double *)c->f[LP_THETA],
                                     (double *)c->f[LP_SGS_T3], (double *)c->f[LP_PSI_M], (T *)c->f[LP_USTAR],
                                     c->d_obuk, g, P));
  HIPCHK(hipGetLastError());
  return LP_OK;
}

// :141-360 theta_all_in_one: filt_da and ddz_uv of theta, scalar_RHS_calc (:474-1110),
// calcbeta (:361-473), the Euler start (:272-274) and step_scalar (:1111-1169).  RHS_T and
// RHS_Tf exchange buffers for :199 (scalar_RHS_calc rewrites every level of RHS_T); on a
// failure they are exchanged back.
static int theta_body(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  const lp_scalar_config &S = cf.scal;
  const lp_grid &gg = c->g;
  Geo g = geo(c);
  const long nv = g.np * g.nz;
  void *th = c->f[LP_THETA], *dTdx = c->f[LP_DTDX], *dTdy = c->f[LP_DTDY], *dTdz = c->f[LP_DTDZ];
  // :160, 175-178 (theta enters the binary32 transform less each level's first value)
  RCHK(do_filt_da(c, th, dTdx, dTdy, gg.nz, nullptr, true, c->tshift));
  RCHK(do_ddz_uv(c, dTdz, th, true));
  const size_t fe = esz_of(c->fdt);
  const double top = S.dTdz_top / S.T_scale * S.z_i;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(g.np), dim3(256), 0, c->s,
                                     (T *)((char *)dTdz + (size_t)(gg.nz - 1) * g.np * fe), g.np, (T)top));
  // :530-575 the surface flux (wt_s through a default REAL, :145, 489) and dsdz(:,:,1)
  const double wts = cf.fp32_casts ? (double)(float)S.wt_s : S.wt_s;
  const double sflux = wts / S.T_scale / S.u_star;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_theta_sfc<T>, dim3(1), dim3(1024), 0, c->s, (const T *)c->f[LP_USTAR],
                                     (const double *)c->d_obuk, (T *)dTdz, g, sflux, cf.vonk, gg.dz));
  HIPCHK(hipGetLastError());
  // :808-879 dealias1 of u, v, w, dsdx, dsdy, dsdz; the products; dealias2
  RowIn vel;
  vel.din = F64;
  const void *src[6] = {c->f[LP_U], c->f[LP_V], c->f[LP_W], dTdx, dTdy, dTdz};
  for (int q = 0; q < 6; ++q) RCHK(to_big(c, src[q], gg.nz, c->tbig[q], q < 3 ? vel : RowIn{}));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_theta_adv<T>, grid1((long)gg.ld_big * gg.ny2 * gg.nz), dim3(256), 0, c->s,
                                     (T *)c->tbig[0], (const T *)c->tbig[1], (const T *)c->tbig[2],
                                     (const T *)c->tbig[3], (const T *)c->tbig[4], (const T *)c->tbig[5], gg.nx2,
                                     gg.ny2, gg.ld_big, gg.nz));
  HIPCHK(hipGetLastError());
  RCHK(from_big(c, c->tbig[0], gg.nz, 1.0 / ((double)gg.nx2 * gg.ny2), c->f[LP_RHS_T]));
  // :889-1060 the SGS flux divergence
  const double inv = 1.0 / S.Pr;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_theta_flux_xy<T>, grid1(nv), dim3(256), 0, c->s, (const T *)c->f[LP_NU_T],
                                     (const T *)dTdx, (const T *)dTdy, (T *)c->tmp[0], (T *)c->tmp[1], g, inv));
  HIPCHK(hipGetLastError());
  const double cst = 1.0 / ((double)gg.nx * gg.ny);
  const int mx[1] = {M_IKX}, my[1] = {M_IKY};
  void *ox[1] = {c->tmp[2]}, *oy[1] = {c->tmp[3]};
  RCHK(spectral_apply(c, c->tmp[0], gg.nz, cst, 1, mx, ox));
  RCHK(spectral_apply(c, c->tmp[1], gg.nz, cst, 1, my, oy));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_theta_flux_z<T>, grid1(nv), dim3(256), 0, c->s, (const T *)c->f[LP_NU_T],
                                     (const T *)dTdz, (T *)c->tmp[0], (double *)c->f[LP_SGS_T3], g, inv, sflux));
  HIPCHK(hipGetLastError());
  RCHK(do_ddz_w(c, c->tmp[1], c->tmp[0]));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_theta_rhs<T>, grid1(nv), dim3(256), 0, c->s, (T *)c->f[LP_RHS_T],
                                     (const T *)c->tmp[2], (const T *)c->tmp[3], (const T *)c->tmp[1], g));
  HIPCHK(hipGetLastError());
  // :246 calcbeta
  hipLaunchKernelGGL(k_level_mean, dim3(gg.nz), dim3(1024), 0, c->s, (const double *)th, g, c->d_tbar);
  const int mode = S.ocean ? S.iboussinesq : 0;
  const double g_hat = S.g * (S.z_i / (S.u_star * S.u_star));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_calcbeta<T>, grid1(nv), dim3(256), 0, c->s, (const double *)th,
                                     (const double *)c->d_tbar, (T *)c->f[LP_BETA_SCAL], g, mode, g_hat,
                                     2.e-4 * S.T_scale));
  HIPCHK(hipGetLastError());
  if (c->jt == S.scal_init && !S.initsc)  // :272-274
    HIPCHK(hipMemcpyAsync(c->f[LP_RHS_TF], c->f[LP_RHS_T], c->bytes[LP_RHS_T], hipMemcpyDeviceToDevice, c->s));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_step_scalar<T>, grid1(g.np * (g.nz - 1)), dim3(256), 0, c->s, (double *)th,
                                     (const T *)c->f[LP_RHS_T], (const T *)c->f[LP_RHS_TF], g, cf.dt));
  hipLaunchKernelGGL(k_scalar_top, grid1(g.np), dim3(256), 0, c->s, (double *)th, g, top * gg.dz,
                     gg.L_z > 1.0 ? 1 : 0);
  HIPCHK(hipGetLastError());
  return LP_OK;
}
static int do_theta(lp_ctx *c) {
  std::swap(c->f[LP_RHS_T], c->f[LP_RHS_TF]);
  const int rc = theta_body(c);
  if (rc != LP_OK) std::swap(c->f[LP_RHS_T], c->f[LP_RHS_TF]);
  return rc;
}

static void swap_F(lp_ctx *c) {
  const int fid[4] = {LP_F_LM, LP_F_MM, LP_F_QN, LP_F_NN};
  for (int q = 0; q < 4; ++q) {
    void *t = c->f[fid[q]];
    c->f[fid[q]] = c->Fi[q];
    c->Fi[q] = (double *)t;
  }
}

static int reset_lag_velocities(lp_ctx *c) {  // interpolag_Sdep.f90:367-369
  const long n = (long)(c->bytes[LP_U_LAG] / esz_of(c->fdt));
  for (int q = 0; q < 3; ++q)
    FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(n), dim3(256), 0, c->s, (T *)c->f[LP_U_LAG + q], n, T(0)));
  HIPCHK(hipGetLastError());
  c->lag_vel_zero = true;
  return LP_OK;
}

// interpolag_Sdep.f90:62-369 on its own: F_LM/MM/QN/NN <- interpolated averages
static int do_interpolag(lp_ctx *c) {
  if (c->cfg.model != 5 || !c->Fi[3]) return set_err(LP_ERR_ARG, "interpolag_Sdep needs model 5");
  Geo g = geo(c);
  const lp_config &cf = c->cfg;
  FDISPATCH(c, T,
            hipLaunchKernelGGL(k_interpolag<T>, grid1(g.np * g.nz), dim3(256), 0, c->s,
                               (const double *)c->f[LP_F_LM], (const double *)c->f[LP_F_MM],
                               (const double *)c->f[LP_F_QN], (const double *)c->f[LP_F_NN], c->Fi[0], c->Fi[1],
                               c->Fi[2], c->Fi[3], (const T *)c->f[LP_U_LAG], (const T *)c->f[LP_V_LAG],
                               (const T *)c->f[LP_W_LAG], g, (double)cf.cs_count, cf.dt, c->g.dx, c->g.dy,
                               c->g.dz));
  HIPCHK(hipGetLastError());
  swap_F(c);
  return reset_lag_velocities(c);
}

// lagrange_Sdep.f90:4-496 (with interpolag_Sdep at :120).  Needs the strain rates in
// c->S and |S| in c->tmp[3] (k_strain).
static int do_convec(lp_ctx *c, bool kept);

// main.f90:419 on the second stream: convec needs u, v, w, their derivatives and the
// wall-stress dudz/dvdz only, so it runs beside the rest of the SGS model, the droplets
// and divstress (its scratch is its own) and is joined before the RHS assembly.  Started
// once per step, after the last column pass of the SGS model.
static int fork_convec(lp_ctx *c) {
  if (!c->fork_pending) return LP_OK;
  c->fork_pending = false;
  HIPCHK(hipEventRecord(c->ev_fork, c->s));
  HIPCHK(hipStreamWaitEvent(c->s2, c->ev_fork, 0));
  std::swap(c->s, c->s2);
  const int rc = do_convec(c, true);  // filt_da left the u, v, w spectra
  std::swap(c->s, c->s2);
  RCHK(rc);
  HIPCHK(hipEventRecord(c->ev_join, c->s2));
  return LP_OK;
}

// use_u: u_lag is zero and this step's accumulation is skipped, so u_lag + u == u and the
// backtrace reads u, v, w directly; u_lag then stays zero, as the reset leaves it
static int do_lagrange(lp_ctx *c, bool use_u = false) {
  const lp_config &cf = c->cfg;
  const lp_grid &gg = c->g;
  Geo g = geo(c);
  const int n1 = c->ncut1, n2 = c->ncut2, nc = n1 > n2 ? n1 : n2;
  const size_t fe = esz_of(c->fdt);
  const double vol = (double)g.np * g.nz;
  // lagrange_Sdep.f90:129-155,192-242: the 21 filter inputs, transformed along x
  LagInputs in;
  in.u = c->f[LP_U]; in.v = c->f[LP_V]; in.w = c->f[LP_W];
  for (int q = 0; q < 6; ++q) in.S[q] = c->S[q];
  in.magS = c->tmp[3];
  prof_begin(c, LP_K_ROW_R2C);
  const long pc = pad16(nc), p1 = pad16(n1), p2 = pad16(n2);
  HIPCHK(lag_r2c(c->fdt, in, g, c->lspec, nc, pc, c->s));
  prof_end(c, LP_K_ROW_R2C, vol * 10.0 * fe + (double)kNumLag * gg.ny * gg.nz * nc * 2 * fe);
  // :157-256 the y part of both test filters, all 21 fields in one pass
  {
    const long ps = pc * gg.ny;
    ColOut o[2] = {{c->lbar, M_G1, p1, p1 * gg.ny, n1}, {c->lhat, M_G2, p2, p2 * gg.ny, n2}};
    // full groups of TX columns in the block-wide pass; one or two trailing columns (nx =
    // 256: the kx = 64 column that G_test keeps by rounding) one column per block
    const int tx = c->fdt == F64 ? 8 : 16;
    const int nmain = (nc % tx <= 2 && nc > tx) ? nc / tx * tx : nc;
    // not in the column-pass class: in a step this launch shares the GPU with convec
    HIPCHK(col_op(c->fdt, gg.ny, gg.ny, 0, 1, 1, c->lspec, pc, ps, nmain, kNumLag * gg.nz, 2, o, pc, ps, c->tab,
                  c->s));
    if (nmain < nc) {
      ColOut ot[2];
      int nt = 0;
      if (nmain < n1) ot[nt++] = o[0];
      if (nmain < n2) ot[nt++] = o[1];
      HIPCHK(col_op_tail(c->fdt, gg.ny, c->lspec, pc, ps, nmain, nc, kNumLag * gg.nz, nt, ot, pc, ps, c->tab,
                         c->s));
    }
  }
  RCHK(fork_convec(c));  // in a step: convec overlaps the rest of the model
  double delta = cf.filter_size * std::pow(gg.dx * gg.dy * gg.dz, 1.0 / 3.0);
  LagParams P;
  P.cst = 2.0 * delta * delta;
  P.tf1_2 = 4.0; P.tf2_2 = 16.0;
  P.opftdelta = cf.opftime * delta;
  P.powcoeff = -1.0 / 8.0;
  P.lagran_dt = cf.dt * (double)cf.cs_count;
  P.bexp = std::log(2.0) / (std::log(4.0) - std::log(2.0));
  P.casts = cf.fp32_casts;
  P.tiny = cf.fp32_casts ? (double)1e-24f : 1e-24;
  P.bclip_min = 1.0 / (2.0 * 4.0);
  P.init_lm = cf.inilag && !c->F_LM_MM_init && (c->jt == cf.cs_count || c->jt == cf.dyn_init);
  P.init_qn = cf.inilag && !c->F_QN_NN_init && (c->jt == cf.cs_count || c->jt == cf.dyn_init);
  const double csd = (double)cf.cs_count;
  LagInterp I{csd, cf.dt * csd / gg.dx, cf.dt * csd / gg.dy, cf.dt * csd / gg.dz};
  const double *Fold[4] = {(const double *)c->f[LP_F_LM], (const double *)c->f[LP_F_MM],
                           (const double *)c->f[LP_F_QN], (const double *)c->f[LP_F_NN]};
  double *Fnew[4] = {c->Fi[0], c->Fi[1], c->Fi[2], c->Fi[3]};
  // :257-326 back along x and contracted, row by row
  prof_begin(c, LP_K_ROW_C2R);
  HIPCHK(lag_c2r_contract(c->fdt, c->lbar, c->lhat, n1, n2, p1, p2, c->lcon, g, P, c->s));
  prof_end(c, LP_K_ROW_C2R, (double)kNumLag * gg.ny * gg.nz * (n1 + n2) * 2 * fe + vol * 4 * fe);
  // :327-434 (+ interpolag_Sdep.f90:62-365): interpolation and relaxation
  prof_begin(c, LP_K_LAG_RELAX);
  const int lv[3] = {use_u ? LP_U : LP_U_LAG, use_u ? LP_V : LP_V_LAG, use_u ? LP_W : LP_W_LAG};
  HIPCHK(lag_relax(c->fdt, c->lcon, Fold, Fnew, c->f[lv[0]], c->f[lv[1]], c->f[lv[2]], c->f[LP_CS_OPT2],
                   c->f[LP_BETA], g, P, I, c->s, use_u));
  prof_end(c, LP_K_LAG_RELAX, vol * (8.0 * 8 + 9.0 * fe));
  swap_F(c);
  if (!use_u) RCHK(reset_lag_velocities(c));
  if (P.init_lm) c->F_LM_MM_init = 1;
  if (P.init_qn) c->F_QN_NN_init = 1;
  return LP_OK;
}

// sgs_stag.f90:188-198, models 2-4 on the generic spectral path (lp_sgs_alt.hpp):
// std_dynamic, scaledep_dynamic, lagrange_Ssim with interpolag_Ssim
static int do_dynamic_alt(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  const lp_grid &gg = c->g;
  Geo g = geo(c);
  const long nv = g.np * g.nz;
  const size_t fe = esz_of(c->fdt);
  const int model = cf.model;
  const double delta = cf.filter_size * std::pow(gg.dx * gg.dy * gg.dz, 1.0 / 3.0);
  const double cst = 2.0 * delta * delta;
  auto vol = [&](void *base, int q) { return (void *)((char *)base + (size_t)q * nv * fe); };
  FDISPATCH(c, T, hipLaunchKernelGGL(k_dyn_inputs<T>, grid1(nv), dim3(256), 0, c->s, (const double *)c->f[LP_U],
                                     (const double *)c->f[LP_V], (const double *)c->f[LP_W], (const T *)c->S[0],
                                     (const T *)c->S[1], (const T *)c->S[2], (const T *)c->S[3],
                                     (const T *)c->S[4], (const T *)c->S[5], (T *)c->dyn_in, g,
                                     c->fdt == F32 ? 1 : 0));
  HIPCHK(hipGetLastError());
  // the test filters (G_test, G_test_test carry the 1/(nx ny) of the transform)
  DynVol B{}, H{};
  for (int q = 0; q < kDynIn; ++q) {
    const void *src = (q >= 9 && q < 15) ? c->S[q - 9] : vol(c->dyn_in, q);
    const bool hat = model == 3 && !(q >= 3 && q < 9);
    const int m[2] = {M_G1, M_G2};
    void *o[2] = {vol(c->dyn_bar, q), hat ? vol(c->dyn_hat, q) : nullptr};
    RCHK(spectral_apply(c, src, gg.nz, 1.0, hat ? 2 : 1, m, o));
    B.p[q] = vol(c->dyn_bar, q);
    if (model == 3) H.p[q] = vol(c->dyn_hat, q);
  }
  if (model == 3) {  // Q_ij: the 2-delta velocities' products, 4-delta filtered (input slots 3-8)
    FDISPATCH(c, T, hipLaunchKernelGGL(k_dyn_qprod<T>, grid1(nv), dim3(256), 0, c->s, (const T *)B.p[0],
                                       (const T *)B.p[1], (const T *)B.p[2], (T *)vol(c->dyn_in, 3), nv));
    HIPCHK(hipGetLastError());
    for (int q = 3; q < 9; ++q) {
      const int m[1] = {M_G2};
      void *o[1] = {vol(c->dyn_hat, q)};
      RCHK(spectral_apply(c, vol(c->dyn_in, q), gg.nz, 1.0, 1, m, o));
    }
    FDISPATCH(c, T, hipLaunchKernelGGL(k_dyn_sums<T>, dim3(gg.nz), dim3(1024), 0, c->s, B, H, g, c->dyn_sums));
    hipLaunchKernelGGL(k_dyn_beta, grid1(gg.nz), dim3(256), 0, c->s, (const double *)c->dyn_sums, c->dyn_beta,
                       gg.nz, delta, (double)gg.nx * gg.ny);
    HIPCHK(hipGetLastError());
  }
  if (model != 4) {
    FDISPATCH(c, T, hipLaunchKernelGGL(k_dyn_cs<T>, dim3(gg.nz), dim3(1024), 0, c->s, B,
                                       (const double *)(model == 3 ? c->dyn_beta : nullptr), g, cst,
                                       (T *)c->f[LP_CS_OPT2], (T *)c->f[LP_BETA]));
    HIPCHK(hipGetLastError());
    return LP_OK;
  }
  // model 4: interpolag_Ssim, then the relaxation of F_LM, F_MM
  const double cs = (double)cf.cs_count;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_interpolag_ssim<T>, grid1(nv), dim3(256), 0, c->s,
                                     (const double *)c->f[LP_F_LM], (const double *)c->f[LP_F_MM], c->Fi[0],
                                     c->Fi[1], (const T *)c->f[LP_U_LAG], (const T *)c->f[LP_V_LAG],
                                     (const T *)c->f[LP_W_LAG], g, cs, cf.dt * cs / gg.dx, cf.dt * cs / gg.dy,
                                     cf.dt * cs / gg.dz));
  HIPCHK(hipGetLastError());
  for (int q = 0; q < 2; ++q) {
    const int fid = q == 0 ? LP_F_LM : LP_F_MM;
    void *t = c->f[fid];
    c->f[fid] = c->Fi[q];
    c->Fi[q] = (double *)t;
  }
  RCHK(reset_lag_velocities(c));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_dyn_contract<T>, grid1(nv), dim3(256), 0, c->s, B, g, cst, (T *)c->tmp[0],
                                     (T *)c->tmp[1]));
  const bool init = cf.inilag && !c->F_LM_MM_init && (c->jt == cf.cs_count || c->jt == cf.dyn_init);
  FDISPATCH(c, T, hipLaunchKernelGGL(k_ssim_relax<T>, grid1(nv), dim3(256), 0, c->s, (const T *)c->tmp[0],
                                     (const T *)c->tmp[1], (double *)c->f[LP_F_LM], (double *)c->f[LP_F_MM],
                                     (T *)c->f[LP_CS_OPT2], (T *)c->f[LP_BETA], g, init ? 1 : 0,
                                     cf.fp32_casts, cf.opftime * delta, cf.dt * cs));
  HIPCHK(hipGetLastError());
  if (init) c->F_LM_MM_init = 1;
  return LP_OK;
}

static int do_sgs_stag(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  const lp_grid &gg = c->g;
  Geo g = geo(c);
  long nv = g.np * g.nz;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_strain<T>, grid1(nv), dim3(256), 0, c->s, (const T *)c->f[LP_DUDX],
                                     (const T *)c->f[LP_DUDY], (const T *)c->f[LP_DUDZ], (const T *)c->f[LP_DVDX],
                                     (const T *)c->f[LP_DVDY], (const T *)c->f[LP_DVDZ], (const T *)c->f[LP_DWDX],
                                     (const T *)c->f[LP_DWDY], (const T *)c->f[LP_DWDZ], (T *)c->S[0],
                                     (T *)c->S[1], (T *)c->S[2], (T *)c->S[3], (T *)c->S[4], (T *)c->S[5],
                                     (T *)c->tmp[3], g));
  HIPCHK(hipGetLastError());
  if (c->frz_nut) {  // frozen coefficient: the base run's Nu_t, no model update
    FDISPATCH(c, T, hipLaunchKernelGGL(k_tau_frozen<T>, grid1(nv), dim3(256), 0, c->s, (const T *)c->S[0],
                                       (const T *)c->S[1], (const T *)c->S[2], (const T *)c->S[3],
                                       (const T *)c->S[4], (const T *)c->S[5], (const T *)c->frz_nut,
                                       (T *)c->f[LP_NU_T], (T *)c->f[LP_TXX], (T *)c->f[LP_TXY], (T *)c->f[LP_TXZ],
                                       (T *)c->f[LP_TYY], (T *)c->f[LP_TYZ], (T *)c->f[LP_TZZ], g));
    HIPCHK(hipGetLastError());
    return LP_OK;
  }
  double delta = cf.filter_size * std::pow(gg.dx * gg.dy * gg.dz, 1.0 / 3.0);
  if (cf.model == 5) {
    const bool init_cs = c->jt == 1 && cf.inilag;  // sgs_stag.f90:178-180
    const bool lag_now = !init_cs && (c->jt >= cf.dyn_init || cf.initu) && (c->jt % cf.cs_count) == 0;
    if (lag_now && c->lag_vel_zero) {
      // sgs_stag.f90:130-134 then :200-201 with u_lag = 0 + u: the model reads u directly
      RCHK(do_lagrange(c, true));
    } else {
      FDISPATCH(c, T, hipLaunchKernelGGL(k_lag_acc<T>, grid1(nv), dim3(256), 0, c->s, (T *)c->f[LP_U_LAG],
                                         (T *)c->f[LP_V_LAG], (T *)c->f[LP_W_LAG], (const double *)c->f[LP_U],
                                         (const double *)c->f[LP_V], (const double *)c->f[LP_W], nv));
      HIPCHK(hipGetLastError());
      c->lag_vel_zero = false;
      if (init_cs) {
        FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(nv), dim3(256), 0, c->s, (T *)c->f[LP_CS_OPT2], nv,
                                           (T)cf.cs2_init));
        HIPCHK(hipGetLastError());
      } else if (lag_now) {
        RCHK(do_lagrange(c));  // sgs_stag.f90:200-201
      }
    }
  } else if (cf.model >= 2) {  // sgs_stag.f90:130-134, 178-198
    if (cf.model == 4) {
      FDISPATCH(c, T, hipLaunchKernelGGL(k_lag_acc<T>, grid1(nv), dim3(256), 0, c->s, (T *)c->f[LP_U_LAG],
                                         (T *)c->f[LP_V_LAG], (T *)c->f[LP_W_LAG], (const double *)c->f[LP_U],
                                         (const double *)c->f[LP_V], (const double *)c->f[LP_W], nv));
      HIPCHK(hipGetLastError());
      c->lag_vel_zero = false;
    }
    if (c->jt == 1 && cf.inilag) {
      FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(nv), dim3(256), 0, c->s, (T *)c->f[LP_CS_OPT2], nv,
                                         (T)cf.cs2_init));
      HIPCHK(hipGetLastError());
    } else if ((c->jt >= cf.dyn_init || cf.initu) && (c->jt % cf.cs_count) == 0) {
      RCHK(do_dynamic_alt(c));
    }
  } else {  // model 1: Cs_opt2 = Co^2 (sgs_stag.f90:148)
    FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(nv), dim3(256), 0, c->s, (T *)c->f[LP_CS_OPT2], nv,
                                       (T)(cf.Co * cf.Co)));
    HIPCHK(hipGetLastError());
  }
  FDISPATCH(c, T, hipLaunchKernelGGL(k_nut_tau<T>, grid1(nv), dim3(256), 0, c->s, (const T *)c->S[0],
                                     (const T *)c->S[1], (const T *)c->S[2], (const T *)c->S[3], (const T *)c->S[4],
                                     (const T *)c->S[5], (const T *)c->f[LP_CS_OPT2], (T *)c->f[LP_NU_T],
                                     (T *)c->f[LP_DISSIP], (T *)c->f[LP_MAGS], (T *)c->f[LP_TXX], (T *)c->f[LP_TXY],
                                     (T *)c->f[LP_TXZ], (T *)c->f[LP_TYY], (T *)c->f[LP_TYZ], (T *)c->f[LP_TZZ], g,
                                     cf.model, delta, cf.Co, cf.nnn, cf.vonk, gg.dz));
  HIPCHK(hipGetLastError());
  return LP_OK;
}

// divstress_uv.f90:31-127 / divstress_w.f90:31-82 from the x-spectra of tx and ty: the
// x- and y-derivatives summed in spectral space (col_div), the z-derivative of tz added
// in the back-transform
static int divstress_from_spectra(lp_ctx *c, void *divt, const void *sx, const void *sy, const void *tz,
                                  int is_w, void *work) {
  const lp_grid &g = c->g;
  const long ps = c->sp * g.ny;
  const size_t fe = esz_of(c->fdt);
  prof_begin(c, LP_K_COL_DIV);
  HIPCHK(col_div(c->fdt, g.ny, sx, sy, c->sp, ps, g.lh, g.nz, work, c->tab, c->s));
  prof_end(c, LP_K_COL_DIV, 3.0 * g.lh * g.ny * g.nz * 2 * fe);
  prof_begin(c, LP_K_ROW_C2R);
  HIPCHK(row_c2r_div(c->fdt, g.nx / 2, work, c->sp, ps, g.lh, g.ny, g.nz, divt, g.ld, plane_el(c), tz,
                     1.0 / g.dz, is_w, c->s));
  prof_end(c, LP_K_ROW_C2R, bytes_c2r(g.nx, g.ny, g.nz, g.lh, fe, fe) + (double)g.ld * g.ny * g.nz * fe);
  return LP_OK;
}

static int do_divstress(lp_ctx *c, void *divt, const void *tx, const void *ty, const void *tz, int is_w) {
  const lp_grid &gg = c->g;
  const double cst = 1.0 / ((double)gg.nx * gg.ny);
  RCHK(r2c_small(c, tx, gg.nz, cst, c->spec[0]));
  RCHK(r2c_small(c, ty, gg.nz, cst, c->spec[1]));
  return divstress_from_spectra(c, divt, c->spec[0], c->spec[1], tz, is_w, c->spec[2]);
}

// main.f90:394-397: the three divergences; txy, txz and tyz are transformed once
static int do_divstress_all(lp_ctx *c) {
  const lp_grid &gg = c->g;
  const double cst = 1.0 / ((double)gg.nx * gg.ny);
  void *sxx = c->spec[0], *sxy = c->spec[1], *syy = c->spec[2], *sxz = c->spec[3], *syz = c->Hs[0];
  void *work = c->Hs[1];
  RCHK(r2c_small(c, c->f[LP_TXX], gg.nz, cst, sxx));
  RCHK(r2c_small(c, c->f[LP_TXY], gg.nz, cst, sxy));
  RCHK(r2c_small(c, c->f[LP_TYY], gg.nz, cst, syy));
  RCHK(r2c_small(c, c->f[LP_TXZ], gg.nz, cst, sxz));
  RCHK(r2c_small(c, c->f[LP_TYZ], gg.nz, cst, syz));
  RCHK(divstress_from_spectra(c, c->f[LP_DIVTX], sxx, sxy, c->f[LP_TXZ], 0, work));
  RCHK(divstress_from_spectra(c, c->f[LP_DIVTY], sxy, syy, c->f[LP_TYZ], 0, work));
  RCHK(divstress_from_spectra(c, c->f[LP_DIVTZ], sxz, syz, c->f[LP_TZZ], 1, work));
  return LP_OK;
}

// dealias1.f90 from the x-spectrum on: the column pass that pads it to ny2 and transforms
// it back along y, into field q of c->cvb; scale multiplies the padded spectrum
static int pad_y(lp_ctx *c, const void *spec, int q, double scale = 1.0) {
  const lp_grid &g = c->g;
  const long ps = c->sp * g.ny, pb = c->cv_pitch * g.ny2;
  const size_t fe = esz_of(c->fdt);
  ColOut o{(char *)c->cvb + (size_t)q * pb * g.nz * 2 * fe, M_ONE};
  o.scale = scale;
  prof_begin(c, LP_K_COL_PAD);
  HIPCHK(col_op(c->fdt, g.ny, g.ny2, 1, 1, 1, spec, c->sp, ps, g.nx / 2, g.nz, 1, &o, c->cv_pitch, pb, c->tab,
                c->s));
  prof_end(c, LP_K_COL_PAD, bytes_col(g.ny, g.ny2, g.nx / 2, g.nz, 1, fe));
  return LP_OK;
}

// kept: c->uvw_spec hold the x-spectra of the current (filtered) u, v, w, as the step's
// filt_da leaves them; otherwise they are transformed here
static int do_convec(lp_ctx *c, bool kept) {
  const lp_grid &gg = c->g;
  Geo g = geo(c);
  const size_t fe = esz_of(c->fdt);
  const double cst = 1.0 / ((double)gg.nx * gg.ny);
  // convec.f90:76-98
  const int vel[3] = {LP_U, LP_V, LP_W};
  RowIn vin;
  vin.din = F64;
  for (int q = 0; q < 3; ++q) {
    if (!kept) RCHK(r2c_small(c, c->f[vel[q]], gg.nz, cst, c->uvw_spec[q], 0, vin));
    if (kept && c->fdt == F32 && q < 2) {
      // filt_da transformed u - s, v - s: the row means (kx = 0 column) get s back
      hipLaunchKernelGGL(k_add_dc<float>, grid1((long)gg.ny * gg.nz), dim3(256), 0, c->s, (cx<float> *)c->uvw_spec[q],
                         c->sp, gg.ny, gg.nz, (const double *)(c->ushift + (size_t)q * gg.nz));
      HIPCHK(hipGetLastError());
    }
    // filt_da's spectrum went through an unnormalised y round trip: ny times the r2c
    RCHK(pad_y(c, c->uvw_spec[q], q, kept ? 1.0 / gg.ny : 1.0));
  }
  // convec.f90:100-133: each vorticity component formed inside its x transform
  const int va[3] = {LP_DWDY, LP_DWDX, LP_DVDX}, vb[3] = {LP_DVDZ, LP_DUDZ, LP_DUDY};
  for (int q = 0; q < 3; ++q) {
    prof_begin(c, LP_K_ROW_R2C);
    HIPCHK(row_r2c_vort(c->fdt, gg.nx / 2, q, c->f[va[q]], c->f[vb[q]], gg.ld, plane_el(c), gg.ny, gg.nz, cst,
                        c->cv_spec[0], c->sp, c->sp * gg.ny, gg.nx / 2, c->s));
    prof_end(c, LP_K_ROW_R2C, bytes_r2c(gg.nx, gg.ny, gg.nz, gg.nx / 2, fe) + (double)vol_el(c) * fe);
    RCHK(pad_y(c, c->cv_spec[0], 3 + q));
  }
  // convec.f90:140-158, 180-221, 241-279 products (normalisation folded in) between the
  // x back- and forward transforms of the 3/2 grid, kx < nx/2 kept (dealias2)
  const double cb = 1.0 / ((double)gg.nx2 * gg.ny2);
  const long pb = c->cv_pitch * gg.ny2;
  prof_begin(c, LP_K_CONVEC);
  HIPCHK(convec_rows(c->fdt, gg.nx2 / 2, c->cvb, c->cvp, c->cv_pitch, gg.ny2, gg.nz, gg.nx / 2, cb, c->s));
  // six padded spectra read, three product spectra written (kx < nx/2)
  prof_end(c, LP_K_CONVEC, (double)(6 * gg.nz + 3 * (gg.nz - 1)) * gg.ny2 * (gg.nx / 2) * 2 * fe);
  // convec.f90:161-170, 227-236, 286-296: unpadd along y, c2r small
  void *outs[3] = {c->f[LP_RHSX], c->f[LP_RHSY], c->f[LP_RHSZ]};
  const long ps = c->sp * gg.ny;
  for (int q = 0; q < 3; ++q) {
    ColOut o{c->cv_spec[1], M_ONE};
    prof_begin(c, LP_K_COL_PAD);
    HIPCHK(col_op(c->fdt, gg.ny2, gg.ny, 2, 1, 1, (const char *)c->cvp + (size_t)q * pb * gg.nz * 2 * fe,
                  c->cv_pitch, pb, gg.nx / 2, gg.nz - 1, 1, &o, c->sp, ps, c->tab, c->s));
    prof_end(c, LP_K_COL_PAD, bytes_col(gg.ny2, gg.ny, gg.nx / 2, gg.nz - 1, 1, fe));
    RCHK(c2r_small(c, c->cv_spec[1], gg.nx / 2, gg.nz - 1, outs[q], c->fdt));
  }
  // convec.f90:306-308 top level BOGUS
  for (int q = 0; q < 3; ++q) {
    FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(g.np), dim3(256), 0, c->s,
                                       (T *)outs[q] + (long)(gg.nz - 1) * g.np, g.np, (T)LP_BOGUS));
    HIPCHK(hipGetLastError());
  }
  return LP_OK;
}

static int do_press(lp_ctx *c) {
  const lp_grid &gg = c->g;
  const lp_config &cf = c->cfg;
  Geo g = geo(c);
  const long nmodes = (long)gg.lh * gg.ny, np = g.np;
  const long ps = c->sp * gg.ny;  // complex elements per spectral plane
  const double tadv1 = 1.5;
  double cst = 1.0 / ((double)gg.nx * gg.ny);
  // :105-122,133-136 wall and lid rows from divtz(1), divtz(nz) (single-plane passes) and
  // the BOGUS top levels of p, dpdx, dpdy (:355-357) on the second stream, beside the H
  // transforms.  A zero pivot sets d_err, which check_err_flag reads and clears, so the
  // flag of every step of an lp_step call is reported.
  auto side = [&]() -> int {
    RCHK(r2c_small(c, c->f[LP_DIVTZ], 1, cst, c->bw));
    ColOut ob{c->bw, M_ONE};
    RCHK(col_small(c, c->bw, 1, 1, 0, 1, &ob));
    const char *divtz_top = (const char *)c->f[LP_DIVTZ] + (size_t)(gg.nz - 1) * np * esz_of(c->fdt);
    RCHK(r2c_small(c, divtz_top, 1, cst, c->tw));
    ColOut ot{c->tw, M_ONE};
    RCHK(col_small(c, c->tw, 1, 1, 0, 1, &ot));
    hipLaunchKernelGGL(k_fill<double>, grid1(np), dim3(256), 0, c->s, (double *)c->f[LP_P] + (long)gg.nz * np, np,
                       LP_BOGUS);
    FDISPATCH(c, T, {
      hipLaunchKernelGGL(k_fill<T>, grid1(np), dim3(256), 0, c->s, (T *)c->f[LP_DPDX] + (long)(gg.nz - 1) * np,
                         np, (T)LP_BOGUS);
      hipLaunchKernelGGL(k_fill<T>, grid1(np), dim3(256), 0, c->s, (T *)c->f[LP_DPDY] + (long)(gg.nz - 1) * np,
                         np, (T)LP_BOGUS);
    });
    HIPCHK(hipGetLastError());
    return LP_OK;
  };
  HIPCHK(hipEventRecord(c->ev_aux_fork, c->s));
  HIPCHK(hipStreamWaitEvent(c->s2, c->ev_aux_fork, 0));
  std::swap(c->s, c->s2);
  c->prof_off = true;
  const int rc_side = side();
  c->prof_off = false;
  std::swap(c->s, c->s2);
  RCHK(rc_side);
  HIPCHK(hipEventRecord(c->ev_aux, c->s2));
  // press_stag_array.f90:75-87,126-131: H = FFT(u*/(1.5 dt))/(nx ny), oddballs zeroed.
  // Binary32 transforms take u* and v* less their value at each level's first point:
  // that moves only the kx = ky = 0 mode of H_x and H_y, which the solve never reads
  // (the mean mode integrates H_z and the wall / lid rows), and it keeps the mean flow's
  // magnitude out of the transforms' rounding.
  const int fid[3] = {LP_U, LP_V, LP_W};
  for (int q = 0; q < 3; ++q) {
    RowIn ri;
    ri.din = F64;
    ri.shift = c->fdt == F32 && q < 2;
    RCHK(r2c_small(c, c->f[fid[q]], gg.nz - 1, cst / tadv1 / cf.dt, c->Hs[q], 0, ri));
    ColOut o{c->Hs[q], M_ONE};
    RCHK(col_small(c, c->Hs[q], gg.nz - 1, 1, 0, 1, &o));
  }
  // :149-329 + tridag_array
  HIPCHK(hipStreamWaitEvent(c->s, c->ev_aux, 0));
  cx<double> *ph = (cx<double> *)c->phat;
  const dim3 grid_modes((unsigned)((nmodes + 63) / 64));  // 64 modes per block: every CU gets some
  prof_begin(c, LP_K_PRESS_TRIDIAG);
  if (c->fdt == F64)
    hipLaunchKernelGGL(k_press_tridiag<double>, grid_modes, dim3(64), 0, c->s, (const cx<double> *)c->Hs[0],
                       (const cx<double> *)c->Hs[1], (const cx<double> *)c->Hs[2], (const cx<double> *)c->bw,
                       (const cx<double> *)c->tw, c->kx, c->ky, ph, c->gam, c->ibet, gg.lh, gg.ny, gg.nz, c->sp,
                       gg.dz, c->d_err);
  else
    hipLaunchKernelGGL(k_press_tridiag<float>, grid_modes, dim3(64), 0, c->s, (const cx<float> *)c->Hs[0],
                       (const cx<float> *)c->Hs[1], (const cx<float> *)c->Hs[2], (const cx<float> *)c->bw,
                       (const cx<float> *)c->tw, c->kx, c->ky, ph, c->pgam, c->ibet, gg.lh, gg.ny, gg.nz, c->sp,
                       gg.dz, c->d_err);
  HIPCHK(hipGetLastError());
  // H_x, H_y, H_z read once, p_hat written once (binary64 complex)
  prof_end(c, LP_K_PRESS_TRIDIAG,
           (double)nmodes * (gg.nz - 1) * 3 * 2 * esz_of(c->fdt) + (double)nmodes * (gg.nz + 1) * 16);
  // :339-352 inverse transforms in binary64: p(0..nz-1), dpdx, dpdy (1..nz-1); dpdx is
  // the back-transform of p's (kx, y) spectrum times i kx (kx is uniform along a column)
  {
    ColOut o[2] = {{c->pspec[0], M_ONE}, {c->pspec[2], M_IKY}};
    HIPCHK(col_op(F64, gg.ny, gg.ny, 0, 0, 1, ph, c->sp, ps, gg.lh, gg.nz, 2, o, c->sp, ps, c->tab, c->s));
  }
  HIPCHK(row_c2r(F64, F64, gg.nx / 2, c->pspec[0], c->sp, ps, gg.lh, gg.ny, gg.nz, c->f[LP_P], gg.ld, np,
                 c->s));
  Dt dto = c->fdt;
  HIPCHK(row_c2r(F64, dto, gg.nx / 2, (const cx<double> *)c->pspec[0] + ps, c->sp, ps, gg.lh, gg.ny,
                 gg.nz - 1, c->f[LP_DPDX], gg.ld, np, c->s, c->kx));
  HIPCHK(row_c2r(F64, dto, gg.nx / 2, (const cx<double> *)c->pspec[2] + ps, c->sp, ps, gg.lh, gg.ny,
                 gg.nz - 1, c->f[LP_DPDY], gg.ld, np, c->s));
  return LP_OK;
}

static DropGeo drop_geo(const lp_ctx *c) {
  DropGeo d;
  d.nx = c->g.nx; d.ny = c->g.ny; d.nz = c->g.nz; d.ld = c->g.ld;
  d.npc = c->cfg.drop.npcon;
  d.np = plane_el(c);
  d.nv = vol_el(c);
  d.fp = (long)c->g.ld * (c->g.ny + 1);
  return d;
}

// scalars_module_crossflow_mod.f90:1894-4472 pollen_RHS_calc2 (+ the optional breakup
// feedback).  prev: the previous RHS after the buffer exchange of pollen_all_in_one2, or
// null when called on its own.
// fuse_step: pollen_all_in_one2 with the breakup source on -- calcbeta_pcon and
// step_pollen2 (and the Euler start's RHS_PCon_f = RHS_PCon) are done by the breakup
// pass; returns 1 in *stepped when it did them
static int do_pollen_rhs(lp_ctx *c, const double *prev, bool fuse_step = false, bool *stepped = nullptr) {
  const lp_droplet_config &D = c->cfg.drop;
  const lp_grid &gg = c->g;
  DropGeo dg = drop_geo(c);
  double *PC = (double *)c->f[LP_PCON], *RH = (double *)c->f[LP_RHS_PCON];
  FDISPATCH(c, T, {
    hipLaunchKernelGGL(k_face_u<T>, dim3((unsigned)(((long)gg.ny * gg.nz + 3) / 4)), dim3(256), 0, c->s,
                       (const double *)c->f[LP_U], (const T *)c->f[LP_DUDX], c->ui0, dg, gg.dx);
    hipLaunchKernelGGL(k_face_v<T>, dim3((unsigned)((gg.nx + 63) / 64), (unsigned)gg.nz), dim3(64 * kFaceSeg), 0,
                       c->s, (const double *)c->f[LP_V], (const T *)c->f[LP_DVDY], c->vi0, dg, gg.dy);
    hipLaunchKernelGGL(k_kc<T>, grid1(dg.nv), dim3(256), 0, c->s, (const T *)c->f[LP_NU_T], (T *)c->f[LP_MAGS],
                       c->kcell, dg, D.Sc);
  });
  HIPCHK(hipGetLastError());
  PollenParams P;
  P.idx = 1.0 / gg.dx; P.idy = 1.0 / gg.dy; P.idz = 1.0 / gg.dz;
  P.idx2 = 1.0 / (gg.dx * gg.dx); P.idy2 = 1.0 / (gg.dy * gg.dy); P.idz2 = 1.0 / (gg.dz * gg.dz);
  P.ghat = D.g / (D.u_star * D.u_star / D.z_i);
  P.psf = D.sfc_flux > 0 ? 0.0 : D.sfc_flux;  // :2287-2467, lbcp = 1
  P.frac = D.src_frac;
  P.frac1 = 1.0 - D.src_frac;
  P.bc_x = D.bc_x; P.bc_y = D.bc_y;
  P.settling = D.settling; P.inertia = D.inertia;
  P.src_on = c->jt_total >= D.src_jt0 && c->jt_total < D.src_jt1;
  P.src_k0 = D.src_k - 1;
  const double nb = (double)dg.nv * dg.npc * 8.0;
  prof_begin(c, LP_K_POLLEN_RHS);
  const dim3 pgrid((unsigned)((dg.nv + 255) / 256), (unsigned)((dg.npc + kPollenBins - 1) / kPollenBins));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_pollen_rhs<T>, pgrid, dim3(256), 0, c->s, PC, RH, prev, c->ui0,
                                     c->vi0, (const double *)c->f[LP_W], (const T *)c->f[LP_DUDT],
                                     (const T *)c->f[LP_DVDT], (const T *)c->f[LP_DWDT], c->kcell, c->d_ws,
                                     c->d_qsrc, c->d_foot, dg, P));
  HIPCHK(hipGetLastError());
  // PCon read once, RHS_PCon written once, plus the bin-independent inputs
  prof_end(c, LP_K_POLLEN_RHS, 2.0 * nb + (double)dg.nv * (8.0 + 3.0 * esz_of(c->fdt)) + 4.0 * dg.fp * gg.nz * 8.0);
  if (D.pbe) {
    prof_begin(c, LP_K_PBE);
    PbeStep St{PC, (double *)c->f[LP_RHS_PCON_F], c->f[LP_BETA_PCON], c->d_vfac, D.active,
               ((c->jt_total == D.pcon_init || c->jt_total == D.src_jt0) && !D.init_pcon) ? 1 : 0, c->cfg.dt};
#define PBE_LAUNCH(NB, STEP)                                                                                      \
  FDISPATCH(c, T, hipLaunchKernelGGL((k_pbe<T, NB, STEP>), grid1(dg.nv), dim3(256), 0, c->s, PC,                \
                                     (double *)c->f[LP_RHS_PBE], RH, (const T *)c->f[LP_DISSIP], c->d_bins,      \
                                     c->d_pbeA, dg, c->kpref, c->ustar3, D.z_i, D.nu_c, D.pbe_feedback, St))
    if (fuse_step) {
      if (dg.npc <= 8) PBE_LAUNCH(8, true);
      else if (dg.npc <= 16) PBE_LAUNCH(16, true);
      else PBE_LAUNCH(32, true);
      if (stepped) *stepped = true;
    } else {
      if (dg.npc <= 8) PBE_LAUNCH(8, false);
      else if (dg.npc <= 16) PBE_LAUNCH(16, false);
      else PBE_LAUNCH(32, false);
    }
#undef PBE_LAUNCH
    HIPCHK(hipGetLastError());
    prof_end(c, LP_K_PBE, (D.pbe_feedback ? 4.0 : 2.0) * nb + (double)dg.nv * esz_of(c->fdt));
  }
  return LP_OK;
}

// :1795-1843 calcbeta_pcon and/or :4481-4508 step_pollen2
static int do_beta_step(lp_ctx *c, int do_beta, int do_step) {
  DropGeo dg = drop_geo(c);
  FDISPATCH(c, T, hipLaunchKernelGGL(k_beta_step<T>, grid1(dg.nv), dim3(256), 0, c->s, (double *)c->f[LP_PCON],
                                     (const double *)c->f[LP_RHS_PCON], (const double *)c->f[LP_RHS_PCON_F],
                                     (T *)c->f[LP_BETA_PCON], c->d_vfac, c->cfg.drop.active, dg, c->cfg.dt, do_beta,
                                     do_step));
  HIPCHK(hipGetLastError());
  return LP_OK;
}

// scalars_module_crossflow_mod.f90:1494-1786 pollen_all_in_one2: RHS_PCon_f = RHS_PCon,
// pollen_RHS_calc2, Euler start, calcbeta_pcon + step_pollen2 (fused)
static int do_pollen(lp_ctx *c) {
  const lp_droplet_config &D = c->cfg.drop;
  // :1521 -- RHS_PCon_f takes the old RHS by exchanging the buffers (k_pollen_rhs
  // rewrites every point of the new RHS_PCon); exchanged back if the RHS fails
  std::swap(c->f[LP_RHS_PCON], c->f[LP_RHS_PCON_F]);
  bool stepped = false;
  const int rc = do_pollen_rhs(c, (const double *)c->f[LP_RHS_PCON_F], true, &stepped);
  if (rc != LP_OK) {
    std::swap(c->f[LP_RHS_PCON], c->f[LP_RHS_PCON_F]);
    return rc;
  }
  if (stepped) return LP_OK;  // the breakup pass did :1660-1665 and the step
  if ((c->jt_total == D.pcon_init || c->jt_total == D.src_jt0) && !D.init_pcon)  // :1660-1665
    HIPCHK(hipMemcpyAsync(c->f[LP_RHS_PCON_F], c->f[LP_RHS_PCON], c->bytes[LP_RHS_PCON], hipMemcpyDeviceToDevice,
                          c->s));
  return do_beta_step(c, 1, 1);
}

// main.f90:864-898: dudt += grad(ke) with ke = energy(); the x/y gradients come from
// filt_da of a copy of ke (only its derivative outputs are used), the z gradient from
// ddz_uv of ke itself
static int do_ke_dudt(lp_ctx *c) {
  const lp_grid &gg = c->g;
  Geo g = geo(c);
  FDISPATCH(c, T, hipLaunchKernelGGL(k_energy<T>, grid1(g.np * g.nz), dim3(256), 0, c->s,
                                     (const double *)c->f[LP_U], (const double *)c->f[LP_V],
                                     (const double *)c->f[LP_W], (T *)c->tmp[0], g));
  HIPCHK(hipGetLastError());
  const int m[2] = {M_IKX, M_IKY};
  void *outs[2] = {c->tmp[1], c->tmp[2]};
  RCHK(spectral_apply(c, c->tmp[0], gg.nz, 1.0 / ((double)gg.nx * gg.ny), 2, m, outs));
  FDISPATCH(c, T, hipLaunchKernelGGL(k_ke_add<T>, grid1(g.np * (g.nz - 1)), dim3(256), 0, c->s, (T *)c->f[LP_DUDT],
                                     (T *)c->f[LP_DVDT], (T *)c->f[LP_DWDT], (const T *)c->tmp[1],
                                     (const T *)c->tmp[2], (const T *)c->tmp[0], g, (T)(1.0 / gg.dz)));
  HIPCHK(hipGetLastError());
  return LP_OK;
}

// reads the flags the kernels raise (zero pivot of the Thomas solves, a non-finite
// velocity written by a step) and clears them for the next call
static int check_err_flag(lp_ctx *c) {
  int h[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(h, c->d_err, sizeof h, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof h, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (h[0]) return set_err(LP_ERR_PIVOT, "tridag_array: zero pivot");
  if (h[1]) return set_err(LP_ERR_NAN, "lp_step: a step wrote a non-finite velocity");
  return LP_OK;
}

// forcing.f90:215-292 boundary values
static int project_bc(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  Geo g = geo(c);
  hipLaunchKernelGGL(k_project_bc<double>, grid1(g.np), dim3(256), 0, c->s, (double *)c->f[LP_U],
                     (double *)c->f[LP_V], (double *)c->f[LP_W], g, cf.use_mean_p_force,
                     cf.lbc_mom == LP_LBC_STRESS_FREE ? 1 : 0);
  HIPCHK(hipGetLastError());
  return LP_OK;
}

static int do_project(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  Geo g = geo(c);
  FDISPATCH(c, T, hipLaunchKernelGGL(k_project<T>, grid1(g.np * (g.nz - 1)), dim3(256), 0, c->s,
                                     (double *)c->f[LP_U], (double *)c->f[LP_V], (double *)c->f[LP_W],
                                     (const T *)c->f[LP_DPDX], (const T *)c->f[LP_DPDY], (const T *)c->f[LP_DPDZ], g,
                                     cf.dt));
  HIPCHK(hipGetLastError());
  return project_bc(c);
}

// main.f90:252-826 after RHS_f = RHS (step_once)
static int step_body(lp_ctx *c) {
  const lp_config &cf = c->cfg;
  const lp_grid &gg = c->g;
  Geo g = geo(c);
  long n1 = g.np * (g.nz - 1);
  const bool theta = cf.scal.s_flag != 0;
  if (theta) RCHK(do_obukhov(c));  // :257, from the unfiltered u(1), v(1)
  // :277-279 (u and v enter the binary32 transforms less a per-level value)
  RCHK(do_filt_da(c, c->f[LP_U], c->f[LP_DUDX], c->f[LP_DUDY], gg.nz, c->uvw_spec[0], true, c->ushift));
  RCHK(do_filt_da(c, c->f[LP_V], c->f[LP_DVDX], c->f[LP_DVDY], gg.nz, c->uvw_spec[1], true, c->ushift + gg.nz));
  RCHK(do_filt_da(c, c->f[LP_W], c->f[LP_DWDX], c->f[LP_DWDY], gg.nz, c->uvw_spec[2], true));
  // :315 wallstress (level 1 of txz, tyz, dudz, dvdz from the test-filtered u(1), v(1):
  // single-plane passes) on the second stream, beside the z-derivatives of the other levels
  HIPCHK(hipEventRecord(c->ev_aux_fork, c->s));
  HIPCHK(hipStreamWaitEvent(c->s2, c->ev_aux_fork, 0));
  std::swap(c->s, c->s2);
  c->prof_off = true;
  const int rc_wall = do_wallstress(c);
  c->prof_off = false;
  std::swap(c->s, c->s2);
  RCHK(rc_wall);
  HIPCHK(hipEventRecord(c->ev_aux, c->s2));
  RCHK(do_ddz_uv(c, c->f[LP_DUDZ], c->f[LP_U], true));  // :284-288 (levels 2..nz-1)
  RCHK(do_ddz_uv(c, c->f[LP_DVDZ], c->f[LP_V], true));
  RCHK(do_ddz_w(c, c->f[LP_DWDZ], c->f[LP_W], true));
  HIPCHK(hipStreamWaitEvent(c->s, c->ev_aux, 0));  // sgs_stag reads level 1
  // convec (:419) needs u, v, w, their derivatives and the wall law's level-1 dudz, dvdz
  // only: it starts on the second stream here and overlaps the whole SGS model
  c->fork_pending = true;
  RCHK(fork_convec(c));
  const int rc_sgs = do_sgs_stag(c);  // :327
  if (rc_sgs != LP_OK) {
    c->fork_pending = false;
    return rc_sgs;
  }
  RCHK(fork_convec(c));  // no Lagrangian step: start it here
  if (theta) {  // :340-344
    if (c->jt >= cf.scal.scal_init) {
      RCHK(do_theta(c));
    } else {
      FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(g.np * g.nz), dim3(256), 0, c->s,
                                         (T *)c->f[LP_BETA_SCAL], g.np * g.nz, T(0)));
      HIPCHK(hipGetLastError());
    }
  }
  const bool drops = cf.drop.npcon > 0;
  if (drops) {
    if (c->jt_total >= cf.drop.pcon_init) {  // :360-372
      RCHK(do_pollen(c));
    } else {  // :352
      FDISPATCH(c, T, hipLaunchKernelGGL(k_fill<T>, grid1(g.np * g.nz), dim3(256), 0, c->s,
                                         (T *)c->f[LP_BETA_PCON], g.np * g.nz, T(0)));
      HIPCHK(hipGetLastError());
    }
  }
  RCHK(do_divstress_all(c));  // :394-397
  HIPCHK(hipStreamWaitEvent(c->s, c->ev_join, 0));
  // :433-439, 458-460, 577-584, 716-728
  const bool first = c->jt == 1 && !cf.initu;
  double force = cf.use_mean_p_force ? cf.mean_p_force : 0.0;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_rhs_ab2<T>, grid1(n1), dim3(256), 0, c->s, (T *)c->f[LP_RHSX],
                                     (T *)c->f[LP_RHSY], (T *)c->f[LP_RHSZ], (const T *)c->f[LP_DIVTX],
                                     (const T *)c->f[LP_DIVTY], (const T *)c->f[LP_DIVTZ],
                                     (T *)(drops ? c->f[LP_DUDT] : nullptr), (T *)c->f[LP_DVDT], (T *)c->f[LP_DWDT],
                                     (const T *)(drops && cf.drop.active ? c->f[LP_BETA_PCON] : nullptr),
                                     (T *)c->f[LP_RHSX_F], (T *)c->f[LP_RHSY_F], (T *)c->f[LP_RHSZ_F], first ? 1 : 0,
                                     (double *)c->f[LP_U], (double *)c->f[LP_V], (double *)c->f[LP_W], n1, cf.dt,
                                     force, (T)cf.mean_p_force,
                                     (const T *)(theta && !cf.scal.passive ? c->f[LP_BETA_SCAL] : nullptr)));
  HIPCHK(hipGetLastError());
  if (first) {  // :577-584 also copies the (BOGUS) top level
    const size_t fe = esz_of(c->fdt), off = (size_t)n1 * fe, pb = (size_t)g.np * fe;
    for (int q = 0; q < 3; ++q)
      HIPCHK(hipMemcpyAsync((char *)c->f[LP_RHSX_F + q] + off, (const char *)c->f[LP_RHSX + q] + off, pb,
                            hipMemcpyDeviceToDevice, c->s));
  }
  RCHK(do_press(c));  // :785
  // :789-803, 826 (forcing.f90:145-292)
  FDISPATCH(c, T, hipLaunchKernelGGL(k_dpdz_project<T>, grid1(g.np * g.nz), dim3(256), 0, c->s,
                                     (const double *)c->f[LP_P], (T *)c->f[LP_DPDZ], (const T *)c->f[LP_DPDX],
                                     (const T *)c->f[LP_DPDY], (T *)c->f[LP_RHSX], (T *)c->f[LP_RHSY],
                                     (T *)c->f[LP_RHSZ], (T *)(drops ? c->f[LP_DUDT] : nullptr), (T *)c->f[LP_DVDT],
                                     (T *)c->f[LP_DWDT], (double *)c->f[LP_U], (double *)c->f[LP_V],
                                     (double *)c->f[LP_W], g, gg.dz, cf.dt, c->d_err + 1));
  HIPCHK(hipGetLastError());
  RCHK(project_bc(c));
  if (drops) RCHK(do_ke_dudt(c));  // :864-898
  return LP_OK;
}

// main.f90:252-826.  :252-254 RHS_f = RHS: the buffers exchange roles (convec rewrites
// every level of RHS); if the step fails before its end they are exchanged back, so that
// RHS_f again holds the last completed step's RHS.
static int step_once(lp_ctx *c) {
  for (int q = 0; q < 3; ++q) std::swap(c->f[LP_RHSX + q], c->f[LP_RHSX_F + q]);
  const int rc = step_body(c);
  if (rc != LP_OK)
    for (int q = 0; q < 3; ++q) std::swap(c->f[LP_RHSX + q], c->f[LP_RHSX_F + q]);
  return rc;
}

static int cfl_of(lp_ctx *c, double *cfl, double *visc);

extern "C" int lp_step(lp_ctx *c, int32_t nsteps) {
  if (!c || nsteps < 0) return set_err(LP_ERR_ARG, "lp_step: bad argument");
  HIPCHK(hipSetDevice(c->dev));
  for (int s = 0; s < nsteps; ++s) {
    RCHK(step_once(c));
    c->jt += 1;
    c->jt_total += 1;
  }
  int rc = check_err_flag(c);
  prof_collect(c);
  if (rc == LP_OK && c->cfl_max > 0.0 && nsteps > 0) {
    double cfl = 0.0, visc = 0.0;
    RCHK(cfl_of(c, &cfl, &visc));
    if (!(cfl <= c->cfl_max))
      return set_err(LP_ERR_CFL, "lp_step: CFL %.4g after step jt_total=%d exceeds the limit %.4g", cfl,
                     c->jt_total - 1, c->cfl_max);
  }
  return rc;
}

extern "C" int lp_set_cfl_limit(lp_ctx *c, double cfl_max) {
  if (!c || !(cfl_max >= 0.0)) return set_err(LP_ERR_ARG, "lp_set_cfl_limit: bad argument");
  c->cfl_max = cfl_max;
  return LP_OK;
}

extern "C" int lp_get_precision(lp_ctx *c, int32_t *precision) {
  if (!c || !precision) return set_err(LP_ERR_ARG, "lp_get_precision: null argument");
  *precision = c->cfg.precision;
  return LP_OK;
}

extern "C" int lp_check_finite(lp_ctx *c) {
  if (!c) return set_err(LP_ERR_ARG, "lp_check_finite: null context");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->s2));
  static const char *names[] = {"u", "v", "w", "RHSx", "RHSy", "RHSz", "RHSx_f", "RHSy_f", "RHSz_f", "p",
                                "F_LM", "F_MM", "F_QN", "F_NN", "Cs_opt2", "PCon", "RHS_PCon", "RHS_PCon_f"};
  static const int ids[] = {LP_U, LP_V, LP_W, LP_RHSX, LP_RHSY, LP_RHSZ, LP_RHSX_F, LP_RHSY_F, LP_RHSZ_F, LP_P,
                            LP_F_LM, LP_F_MM, LP_F_QN, LP_F_NN, LP_CS_OPT2, LP_PCON, LP_RHS_PCON, LP_RHS_PCON_F};
  const int nchk = (int)(sizeof ids / sizeof ids[0]) - (c->cfg.drop.npcon > 0 ? 0 : 3);
  std::vector<int> flags(nchk, 0);
  for (int q = 0; q < nchk; ++q) {
    HIPCHK(hipMemsetAsync(c->d_err + 2, 0, sizeof(int), c->s));
    const int id = ids[q];
    const long n = (long)(c->bytes[id] / c->esz[id]);
    if (c->esz[id] == 8)
      hipLaunchKernelGGL(k_nonfinite<double>, dim3(c->nred), dim3(256), 0, c->s, (const double *)c->f[id], n,
                         c->d_err + 2);
    else
      hipLaunchKernelGGL(k_nonfinite<float>, dim3(c->nred), dim3(256), 0, c->s, (const float *)c->f[id], n,
                         c->d_err + 2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&flags[q], c->d_err + 2, sizeof(int), hipMemcpyDeviceToHost, c->s));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  for (int q = 0; q < nchk; ++q)
    if (flags[q]) return set_err(LP_ERR_NAN, "lp_check_finite: %s holds a non-finite value", names[q]);
  return LP_OK;
}

extern "C" int lp_prof_enable(lp_ctx *c, int32_t kernel_class) {
  if (!c || kernel_class < -1 || kernel_class >= LP_K_NCLASSES)
    return set_err(LP_ERR_ARG, "lp_prof_enable: bad argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->s));
  prof_collect(c);
  c->prof_class = kernel_class;
  c->prof_ms = 0.0;
  c->prof_bytes = 0.0;
  c->prof_launches = 0;
  return LP_OK;
}

extern "C" int lp_prof_read(lp_ctx *c, double *total_ms, int64_t *launches, double *bytes) {
  if (!c) return set_err(LP_ERR_ARG, "lp_prof_read: null context");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->s));
  prof_collect(c);
  if (total_ms) *total_ms = c->prof_ms;
  if (launches) *launches = c->prof_launches;
  if (bytes) *bytes = c->prof_bytes;
  return LP_OK;
}

// ------------------------------------------------------------------------------------
// state-routine entry points
// ------------------------------------------------------------------------------------
#define STATE_ENTRY(name, body)                                   \
  extern "C" int name(lp_ctx *c) {                                \
    if (!c) return set_err(LP_ERR_ARG, #name ": null context");   \
    HIPCHK(hipSetDevice(c->dev));                                 \
    RCHK(body);                                                   \
    HIPCHK(hipStreamSynchronize(c->s));                           \
    return LP_OK;                                                 \
  }
STATE_ENTRY(lp_wallstress, do_wallstress(c))
STATE_ENTRY(lp_sgs_stag, do_sgs_stag(c))
STATE_ENTRY(lp_interpolag_sdep, do_interpolag(c))
STATE_ENTRY(lp_convec, do_convec(c, false))
STATE_ENTRY(lp_project, do_project(c))

extern "C" int lp_lagrange_sdep(lp_ctx *c) {
  if (!c) return set_err(LP_ERR_ARG, "lp_lagrange_sdep: null context");
  if (c->cfg.model != 5) return set_err(LP_ERR_ARG, "lp_lagrange_sdep needs model 5");
  HIPCHK(hipSetDevice(c->dev));
  // the strain rates come from the derivative fields, as sgs_stag computes them first
  Geo g = geo(c);
  long nv = g.np * g.nz;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_strain<T>, grid1(nv), dim3(256), 0, c->s, (const T *)c->f[LP_DUDX],
                                     (const T *)c->f[LP_DUDY], (const T *)c->f[LP_DUDZ], (const T *)c->f[LP_DVDX],
                                     (const T *)c->f[LP_DVDY], (const T *)c->f[LP_DVDZ], (const T *)c->f[LP_DWDX],
                                     (const T *)c->f[LP_DWDY], (const T *)c->f[LP_DWDZ], (T *)c->S[0],
                                     (T *)c->S[1], (T *)c->S[2], (T *)c->S[3], (T *)c->S[4], (T *)c->S[5],
                                     (T *)c->tmp[3], g));
  HIPCHK(hipGetLastError());
  RCHK(do_lagrange(c));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}

#define DROP_ENTRY(name, body)                                                      \
  extern "C" int name(lp_ctx *c) {                                                  \
    if (!c) return set_err(LP_ERR_ARG, #name ": null context");                     \
    if (c->cfg.drop.npcon < 1) return set_err(LP_ERR_ARG, #name " needs npcon > 0"); \
    HIPCHK(hipSetDevice(c->dev));                                                   \
    RCHK(body);                                                                     \
    HIPCHK(hipStreamSynchronize(c->s));                                             \
    return LP_OK;                                                                   \
  }
DROP_ENTRY(lp_pollen_all_in_one2, do_pollen(c))

#define THETA_ENTRY(name, body)                                                        \
  extern "C" int name(lp_ctx *c) {                                                     \
    if (!c) return set_err(LP_ERR_ARG, #name ": null context");                        \
    if (!c->cfg.scal.s_flag) return set_err(LP_ERR_ARG, #name " needs scal.s_flag"); \
    HIPCHK(hipSetDevice(c->dev));                                                      \
    RCHK(body);                                                                        \
    HIPCHK(hipStreamSynchronize(c->s));                                                \
    return LP_OK;                                                                      \
  }
THETA_ENTRY(lp_obukhov, do_obukhov(c))
THETA_ENTRY(lp_theta_all_in_one, do_theta(c))
extern "C" int lp_obukhov_get(lp_ctx *c, double *out) {
  if (!c || !out) return set_err(LP_ERR_ARG, "lp_obukhov_get: null argument");
  if (!c->cfg.scal.s_flag) return set_err(LP_ERR_ARG, "lp_obukhov_get needs scal.s_flag");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipMemcpyAsync(out, c->d_obuk, sizeof(double) * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}
DROP_ENTRY(lp_pollen_rhs_calc2, do_pollen_rhs(c, nullptr))
DROP_ENTRY(lp_calcbeta_pcon, do_beta_step(c, 1, 0))
DROP_ENTRY(lp_step_pollen2, do_beta_step(c, 0, 1))

extern "C" int lp_press_stag_array(lp_ctx *c) {
  if (!c) return set_err(LP_ERR_ARG, "lp_press_stag_array: null context");
  HIPCHK(hipSetDevice(c->dev));
  RCHK(do_press(c));
  return check_err_flag(c);
}

extern "C" int lp_timestep_conditions(lp_ctx *c, double *cfl, double *visc_stab) {
  if (!c || !cfl || !visc_stab) return set_err(LP_ERR_ARG, "lp_timestep_conditions: null argument");
  HIPCHK(hipSetDevice(c->dev));
  return cfl_of(c, cfl, visc_stab);
}

// scalars_module2.f90:1336-1379
static int cfl_of(lp_ctx *c, double *cfl, double *visc_stab) {
  Geo g = geo(c);
  FDISPATCH(c, T, hipLaunchKernelGGL(k_reduce_diag<T>, dim3(c->nred), dim3(256), 0, c->s, (const double *)c->f[LP_U],
                                     (const double *)c->f[LP_V], (const double *)c->f[LP_W], (const T *)c->f[LP_NU_T],
                                     (const T *)c->f[LP_DUDX], (const T *)c->f[LP_DVDY], (const T *)c->f[LP_DWDZ], g,
                                     c->d_part));
  HIPCHK(hipGetLastError());
  std::vector<double> h(3 * (size_t)c->nred);
  HIPCHK(hipMemcpyAsync(h.data(), c->d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  double mu = 0.0, mn = -INFINITY;
  for (int b = 0; b < c->nred; ++b) { mu = std::fmax(mu, h[3 * b]); mn = std::fmax(mn, h[3 * b + 1]); }
  double delta = std::fmin(c->g.dx, std::fmin(c->g.dy, c->g.dz));
  *cfl = std::sqrt(mu) * c->cfg.dt / delta;
  *visc_stab = c->cfg.dt * mn / (delta * delta);
  return LP_OK;
}

extern "C" int lp_rmsdiv(lp_ctx *c, double *rms) {
  if (!c || !rms) return set_err(LP_ERR_ARG, "lp_rmsdiv: null argument");
  HIPCHK(hipSetDevice(c->dev));
  Geo g = geo(c);
  FDISPATCH(c, T, hipLaunchKernelGGL(k_reduce_diag<T>, dim3(c->nred), dim3(256), 0, c->s, (const double *)c->f[LP_U],
                                     (const double *)c->f[LP_V], (const double *)c->f[LP_W], (const T *)c->f[LP_NU_T],
                                     (const T *)c->f[LP_DUDX], (const T *)c->f[LP_DVDY], (const T *)c->f[LP_DWDZ], g,
                                     c->d_part));
  HIPCHK(hipGetLastError());
  std::vector<double> h(3 * (size_t)c->nred);
  HIPCHK(hipMemcpyAsync(h.data(), c->d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  double s = 0.0;
  for (int b = 0; b < c->nred; ++b) s += h[3 * b + 2];
  *rms = s / ((double)c->g.nx * c->g.ny * (c->g.nz - 1));
  return LP_OK;
}

extern "C" int lp_plane_stats(lp_ctx *c, double *out) {
  if (!c || !out) return set_err(LP_ERR_ARG, "lp_plane_stats: null argument");
  HIPCHK(hipSetDevice(c->dev));
  Geo g = geo(c);
  const int nl = g.nz - 1;
  FDISPATCH(c, T, {
    hipLaunchKernelGGL((k_plane_stats<T, 0>), dim3(nl), dim3(256), 0, c->s, (const double *)c->f[LP_U],
                       (const double *)c->f[LP_V], (const double *)c->f[LP_W], (const T *)c->f[LP_CS_OPT2],
                       (const T *)c->f[LP_NU_T], (const T *)c->f[LP_TXZ], g, c->d_stats);
    hipLaunchKernelGGL((k_plane_stats<T, 1>), dim3(nl), dim3(256), 0, c->s, (const double *)c->f[LP_U],
                       (const double *)c->f[LP_V], (const double *)c->f[LP_W], (const T *)c->f[LP_CS_OPT2],
                       (const T *)c->f[LP_NU_T], (const T *)c->f[LP_TXZ], g, c->d_stats);
  });
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->d_stats, sizeof(double) * 10 * (size_t)nl, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}

// energy.f90:1-47: the mean kinetic energy of the uv nodes of levels 1..nz-1 (the level
// sums on the device in binary64, added level by level on the host)
extern "C" int lp_energy(lp_ctx *c, double *kea) {
  if (!c || !kea) return set_err(LP_ERR_ARG, "lp_energy: null argument");
  HIPCHK(hipSetDevice(c->dev));
  Geo g = geo(c);
  const int nl = g.nz - 1;
  hipLaunchKernelGGL(k_level_ke, dim3(nl), dim3(256), 0, c->s, (const double *)c->f[LP_U],
                     (const double *)c->f[LP_V], (const double *)c->f[LP_W], g, c->d_stats);
  HIPCHK(hipGetLastError());
  std::vector<double> lev(nl);
  HIPCHK(hipMemcpyAsync(lev.data(), c->d_stats, sizeof(double) * nl, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  double s = 0.0;
  for (int k = 0; k < nl; ++k) s += lev[k];
  *kea = s / ((double)g.nx * g.ny * nl);
  return LP_OK;
}

// io_crossflow_mod.f90:2452-2642 tavg_init (the accumulators start at zero) and
// :2647-2799 tavg_compute
static int tavg_alloc(lp_ctx *c) {
  const long nv = vol_el(c);
  const int npc = c->cfg.drop.npcon;
  const int nq = npc > 0 ? 30 + 2 * npc : 28;
  if (!c->tavg) {
    RCHK(dalloc((void **)&c->tavg, sizeof(double) * (size_t)nq * nv));
    c->tavg_nq = nq;
  }
  hipLaunchKernelGGL(k_fill<double>, grid1((long)nq * nv), dim3(256), 0, c->s, c->tavg, (long)nq * nv, 0.0);
  HIPCHK(hipGetLastError());
  c->tavg_time = 0.0;
  return LP_OK;
}
extern "C" int lp_tavg_reset(lp_ctx *c) {
  if (!c) return set_err(LP_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  RCHK(tavg_alloc(c));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}
extern "C" int lp_tavg_compute(lp_ctx *c, double tavg_dt) {
  if (!c) return set_err(LP_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  if (!c->tavg) RCHK(tavg_alloc(c));
  Geo g = geo(c);
  const int npc = c->cfg.drop.npcon;
  FDISPATCH(c, T, hipLaunchKernelGGL(k_tavg<T>, grid1(g.np * g.nz), dim3(256), 0, c->s, c->tavg,
                                     (const double *)c->f[LP_U], (const double *)c->f[LP_V],
                                     (const double *)c->f[LP_W], (const double *)c->f[LP_P],
                                     (const T *)c->f[LP_TXX], (const T *)c->f[LP_TYY], (const T *)c->f[LP_TZZ],
                                     (const T *)c->f[LP_TXY], (const T *)c->f[LP_TXZ], (const T *)c->f[LP_TYZ],
                                     (const T *)c->f[LP_DUDX], (const T *)c->f[LP_DUDY], (const T *)c->f[LP_DUDZ],
                                     (const T *)c->f[LP_DVDX], (const T *)c->f[LP_DVDY], (const T *)c->f[LP_DVDZ],
                                     (const T *)c->f[LP_DWDX], (const T *)c->f[LP_DWDY], (const T *)c->f[LP_DWDZ],
                                     (const double *)(npc > 0 ? c->f[LP_PCON] : nullptr),
                                     (const double *)c->d_pvol, npc, g, tavg_dt));
  HIPCHK(hipGetLastError());
  c->tavg_time += tavg_dt;  // :2795 tavg_total_time
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}
extern "C" int lp_tavg_get(lp_ctx *c, int32_t q, void *out, size_t bytes, double *total_time) {
  if (!c || !out) return set_err(LP_ERR_ARG, "lp_tavg_get: null argument");
  if (!c->tavg) return set_err(LP_ERR_ARG, "lp_tavg_get: no averages yet (lp_tavg_reset / lp_tavg_compute)");
  const size_t vb = sizeof(double) * (size_t)vol_el(c);
  if (q < 0 || q >= c->tavg_nq) return set_err(LP_ERR_ARG, "lp_tavg_get: quantity %d of %d", q, c->tavg_nq);
  if (bytes != vb) return set_err(LP_ERR_ARG, "lp_tavg_get: %zu bytes, expected %zu", bytes, vb);
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipMemcpyAsync(out, c->tavg + (size_t)q * vol_el(c), vb, hipMemcpyDefault, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (total_time) *total_time = c->tavg_time;
  return LP_OK;
}

// io.f90:1643-1679 on levels 1..nz-1: the rows of the volume (bin `bin` of a per-bin
// volume) go through the x transform, scaled by 1/nx, in the field's own type, then
// k_spectrum forms the y-averaged power per wavenumber.
extern "C" int lp_spectrum(lp_ctx *c, int32_t field, int32_t bin, double *out) {
  if (!c || !out) return set_err(LP_ERR_ARG, "lp_spectrum: null argument");
  const long nv = vol_el(c);
  if (field < 0 || field >= LP_NFIELDS || field == LP_P || field == LP_USTAR || bin < 0 ||
      c->bytes[field] < (size_t)(bin + 1) * nv * c->esz[field])
    return set_err(LP_ERR_ARG, "lp_spectrum: field %d bin %d is not a volume of the context", field, bin);
  HIPCHK(hipSetDevice(c->dev));
  const lp_grid &g = c->g;
  const int nl = g.nz - 1, nh = g.nx / 2;
  const bool f64 = c->esz[field] == 8;
  const Dt dt = f64 ? F64 : c->fdt;
  void *spec = f64 ? c->pspec[0] : c->spec[0];  // (nz+1) resp. nz levels of (sp, ny) complex
  const char *in = (const char *)c->f[field] + (size_t)bin * nv * c->esz[field];
  HIPCHK(row_r2c(dt, nh, in, g.ld, plane_el(c), g.ny, nl, 1.0 / g.nx, spec, c->sp, c->sp * g.ny, nh, c->s));
  const dim3 grid((nh + 63) / 64, nl);
  if (f64)
    hipLaunchKernelGGL(k_spectrum<double>, grid, dim3(64), 0, c->s, (const double *)spec, c->sp, g.ny, nh,
                       c->d_espec);
  else
    hipLaunchKernelGGL(k_spectrum<float>, grid, dim3(64), 0, c->s, (const float *)spec, c->sp, g.ny, nh,
                       c->d_espec);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->d_espec, sizeof(double) * (size_t)nh * nl, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}

// ------------------------------------------------------------------------------------
// explicit-array entry points: host or device pointers
// ------------------------------------------------------------------------------------
namespace {
bool is_device_ptr(const void *p) {
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof a);
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// A view of a caller array on the device: device pointers are used in place, host
// pointers are copied to a temporary device buffer (and back when `out`).
struct Arg {
  void *host = nullptr;
  void *dev = nullptr;
  size_t bytes = 0;
  bool owned = false, out = false;
  int bind(lp_ctx *c, const void *p, size_t b, bool in, bool out_) {
    bytes = b;
    out = out_;
    if (!p) return set_err(LP_ERR_ARG, "null array argument");
    if (is_device_ptr(p)) {
      dev = const_cast<void *>(p);
      return LP_OK;
    }
    host = const_cast<void *>(p);
    owned = true;
    hipError_t e = hipMalloc(&dev, b);
    if (e != hipSuccess) return set_err(LP_ERR_NOMEM, "staging hipMalloc(%zu) failed", b);
    if (in) {
      e = hipMemcpyAsync(dev, p, b, hipMemcpyHostToDevice, c->s);
      if (e != hipSuccess) return set_err(LP_ERR_HIP, "staging copy failed: %s", hipGetErrorString(e));
    } else {
      e = hipMemsetAsync(dev, 0, b, c->s);
      if (e != hipSuccess) return set_err(LP_ERR_HIP, "staging memset failed");
    }
    return LP_OK;
  }
  int finish(lp_ctx *c) {
    if (owned && out) {
      hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->s);
      if (e != hipSuccess) return set_err(LP_ERR_HIP, "staging copy-back failed");
    }
    return LP_OK;
  }
  ~Arg() {
    if (owned && dev) (void)hipFree(dev);
  }
};

int finish_all(lp_ctx *c, std::initializer_list<Arg *> args) {
  for (Arg *a : args) RCHK(a->finish(c));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}
}  // namespace

static size_t vbytes(lp_ctx *c, int nlev) { return (size_t)plane_el(c) * nlev * esz_of(c->fdt); }

#define ENTRY_PROLOGUE(c)                                        \
  if (!c) return set_err(LP_ERR_ARG, "null context");            \
  HIPCHK(hipSetDevice(c->dev));

extern "C" int lp_init_wavenumber(lp_ctx *c, double *kx, double *ky, double *k2) {
  ENTRY_PROLOGUE(c);
  size_t b = sizeof(double) * (size_t)c->g.lh * c->g.ny;
  Arg a, bb, cc;
  RCHK(a.bind(c, kx, b, false, true));
  RCHK(bb.bind(c, ky, b, false, true));
  RCHK(cc.bind(c, k2, b, false, true));
  HIPCHK(hipMemcpyAsync(a.dev, c->kx, b, hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(bb.dev, c->ky, b, hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(cc.dev, c->k2, b, hipMemcpyDeviceToDevice, c->s));
  return finish_all(c, {&a, &bb, &cc});
}

extern "C" int lp_test_filter_init(lp_ctx *c, double alpha, double *G) {
  ENTRY_PROLOGUE(c);
  std::vector<double> kx, ky, k2, Gh;
  host_wavenumbers(c->g, kx, ky, k2);
  host_filter(c->g, k2, alpha, Gh);
  size_t b = sizeof(double) * Gh.size();
  Arg a;
  RCHK(a.bind(c, G, b, false, true));
  HIPCHK(hipMemcpyAsync(a.dev, Gh.data(), b, hipMemcpyHostToDevice, c->s));
  return finish_all(c, {&a});
}

extern "C" int lp_filt_da(lp_ctx *c, void *f, void *dfdx, void *dfdy, int32_t nlev) {
  ENTRY_PROLOGUE(c);
  if (nlev < 1) return set_err(LP_ERR_ARG, "nlev < 1");
  size_t b = vbytes(c, nlev);
  Arg a, x, y;
  RCHK(a.bind(c, f, b, true, true));
  RCHK(x.bind(c, dfdx, b, false, true));
  RCHK(y.bind(c, dfdy, b, false, true));
  RCHK(do_filt_da(c, a.dev, x.dev, y.dev, nlev));
  return finish_all(c, {&a, &x, &y});
}

static int deriv_entry(lp_ctx *c, void *dfdx, void *dfdy, const void *f, int nlev) {
  ENTRY_PROLOGUE(c);
  if (nlev < 1) return set_err(LP_ERR_ARG, "nlev < 1");
  size_t b = vbytes(c, nlev);
  Arg a, x, y;
  RCHK(a.bind(c, f, b, true, false));
  int m[2];
  void *outs[2];
  int n = 0;
  if (dfdx) { RCHK(x.bind(c, dfdx, b, false, true)); m[n] = M_IKX; outs[n++] = x.dev; }
  if (dfdy) { RCHK(y.bind(c, dfdy, b, false, true)); m[n] = M_IKY; outs[n++] = y.dev; }
  double cst = 1.0 / ((double)c->g.nx * c->g.ny);
  RCHK(spectral_apply(c, a.dev, nlev, cst, n, m, outs));
  return finish_all(c, {&a, &x, &y});
}
extern "C" int lp_ddx(lp_ctx *c, void *dfdx, const void *f, int32_t nlev) {
  if (!dfdx) return set_err(LP_ERR_ARG, "null array argument");
  return deriv_entry(c, dfdx, nullptr, f, nlev);
}
extern "C" int lp_ddy(lp_ctx *c, void *dfdy, const void *f, int32_t nlev) {
  if (!dfdy) return set_err(LP_ERR_ARG, "null array argument");
  return deriv_entry(c, nullptr, dfdy, f, nlev);
}
extern "C" int lp_ddxy(lp_ctx *c, void *dfdx, void *dfdy, const void *f, int32_t nlev) {
  if (!dfdx || !dfdy) return set_err(LP_ERR_ARG, "null array argument");
  return deriv_entry(c, dfdx, dfdy, f, nlev);
}

extern "C" int lp_ddz_uv(lp_ctx *c, void *dfdz, const void *f) {
  ENTRY_PROLOGUE(c);
  size_t b = vbytes(c, c->g.nz);
  Arg a, d;
  RCHK(a.bind(c, f, b, true, false));
  RCHK(d.bind(c, dfdz, b, true, true));  // level 1 is untouched: keep the caller's values
  RCHK(do_ddz_uv(c, d.dev, a.dev));
  return finish_all(c, {&a, &d});
}
extern "C" int lp_ddz_w(lp_ctx *c, void *dfdz, const void *f) {
  ENTRY_PROLOGUE(c);
  size_t b = vbytes(c, c->g.nz);
  Arg a, d;
  RCHK(a.bind(c, f, b, true, false));
  RCHK(d.bind(c, dfdz, b, true, true));
  RCHK(do_ddz_w(c, d.dev, a.dev));
  return finish_all(c, {&a, &d});
}

extern "C" int lp_padd(lp_ctx *c, void *u_big, const void *u, int32_t nplanes) {
  ENTRY_PROLOGUE(c);
  const lp_grid &g = c->g;
  size_t fe = esz_of(c->fdt);
  long ps = (long)g.lh * g.ny, pb = (long)g.lh_big * g.ny2;
  Arg a, o;
  RCHK(a.bind(c, u, (size_t)ps * nplanes * 2 * fe, true, false));
  RCHK(o.bind(c, u_big, (size_t)pb * nplanes * 2 * fe, false, true));
  HIPCHK(hipMemsetAsync(o.dev, 0, o.bytes, c->s));  // padd.f90:12
  ColOut co{o.dev, M_ONE};
  HIPCHK(col_op(c->fdt, g.ny, g.ny2, 1, 0, 0, a.dev, g.lh, ps, g.nx / 2, nplanes, 1, &co, g.lh_big, pb, c->tab,
                c->s));
  return finish_all(c, {&a, &o});
}

extern "C" int lp_unpadd(lp_ctx *c, void *cc, const void *cc_big, int32_t nplanes) {
  ENTRY_PROLOGUE(c);
  const lp_grid &g = c->g;
  size_t fe = esz_of(c->fdt);
  long ps = (long)g.lh * g.ny, pb = (long)g.lh_big * g.ny2;
  Arg a, o;
  RCHK(a.bind(c, cc_big, (size_t)pb * nplanes * 2 * fe, true, false));
  RCHK(o.bind(c, cc, (size_t)ps * nplanes * 2 * fe, true, true));
  // unpadd.f90:15-16 zero the Nyquist column and row; the column map writes the rest
  // (kx < nx/2).  The kx = nx/2 column is cleared by a zero-column map.
  ColOut co{o.dev, M_ONE};
  HIPCHK(col_op(c->fdt, g.ny2, g.ny, 2, 0, 0, a.dev, g.lh_big, pb, g.nx / 2, nplanes, 1, &co, g.lh, ps, c->tab,
                c->s));
  // Nyquist column jx = nx/2: write zeros via a padmode-0 map of ncols = lh that only
  // touches the last column: simplest is a strided fill kernel.
  for (int p = 0; p < nplanes; ++p) {
    HIPCHK(hipMemset2DAsync((char *)o.dev + ((size_t)p * ps + g.nx / 2) * 2 * fe, (size_t)g.lh * 2 * fe, 0,
                            2 * fe, g.ny, c->s));
  }
  return finish_all(c, {&a, &o});
}

extern "C" int lp_dealias1(lp_ctx *c, const void *u, void *u_big, int32_t nlev) {
  ENTRY_PROLOGUE(c);
  const lp_grid &g = c->g;
  Arg a, o;
  RCHK(a.bind(c, u, vbytes(c, nlev), true, false));
  RCHK(o.bind(c, u_big, (size_t)g.ld_big * g.ny2 * nlev * esz_of(c->fdt), false, true));
  RCHK(to_big(c, a.dev, nlev, o.dev));
  return finish_all(c, {&a, &o});
}

extern "C" int lp_dealias2(lp_ctx *c, void *u, void *u_big, int32_t nlev) {
  ENTRY_PROLOGUE(c);
  const lp_grid &g = c->g;
  Arg a, o;
  RCHK(a.bind(c, u_big, (size_t)g.ld_big * g.ny2 * nlev * esz_of(c->fdt), true, false));
  RCHK(o.bind(c, u, vbytes(c, nlev), false, true));
  RCHK(from_big(c, a.dev, nlev, 1.0 / ((double)g.nx2 * g.ny2), o.dev));
  return finish_all(c, {&a, &o});
}

extern "C" int lp_test_filter(lp_ctx *c, void *f, const double *G, int32_t nplanes) {
  ENTRY_PROLOGUE(c);
  const lp_grid &g = c->g;
  size_t gb = sizeof(double) * (size_t)g.lh * g.ny;
  Arg a, gg;
  RCHK(a.bind(c, f, vbytes(c, nplanes), true, true));
  RCHK(gg.bind(c, G, gb, true, false));
  SpecTables t = c->tab;
  t.G1 = (const double *)gg.dev;
  RCHK(r2c_small(c, a.dev, nplanes, 1.0, c->spec[0]));
  ColOut o{c->spec[0], M_G1};
  const long ps = c->sp * g.ny;
  HIPCHK(col_op(c->fdt, g.ny, g.ny, 0, 1, 1, c->spec[0], c->sp, ps, g.lh, nplanes, 1, &o, c->sp, ps, t, c->s));
  RCHK(c2r_small(c, c->spec[0], g.lh, nplanes, a.dev, c->fdt));
  return finish_all(c, {&a, &gg});
}

extern "C" int lp_divstress_uv(lp_ctx *c, void *divt, const void *tx, const void *ty, const void *tz) {
  ENTRY_PROLOGUE(c);
  size_t b = vbytes(c, c->g.nz);
  Arg d, a, bb, cc;
  RCHK(d.bind(c, divt, b, false, true));
  RCHK(a.bind(c, tx, b, true, false));
  RCHK(bb.bind(c, ty, b, true, false));
  RCHK(cc.bind(c, tz, b, true, false));
  RCHK(do_divstress(c, d.dev, a.dev, bb.dev, cc.dev, 0));
  return finish_all(c, {&d, &a, &bb, &cc});
}

extern "C" int lp_divstress_w(lp_ctx *c, void *divt, const void *tx, const void *ty, const void *tz) {
  ENTRY_PROLOGUE(c);
  size_t b = vbytes(c, c->g.nz);
  Arg d, a, bb, cc;
  RCHK(d.bind(c, divt, b, false, true));
  RCHK(a.bind(c, tx, b, true, false));
  RCHK(bb.bind(c, ty, b, true, false));
  RCHK(cc.bind(c, tz, b, true, false));
  RCHK(do_divstress(c, d.dev, a.dev, bb.dev, cc.dev, 1));
  return finish_all(c, {&d, &a, &bb, &cc});
}

extern "C" int lp_tridag_array(lp_ctx *c, const double *a, const double *b, const double *cc, const double *r,
                               double *u) {
  ENTRY_PROLOGUE(c);
  const lp_grid &g = c->g;
  long nh = (long)g.lh * g.ny;
  int n = g.nz + 1;
  size_t rb = sizeof(double) * nh * n;
  Arg A, B, Cc, R, U;
  RCHK(A.bind(c, a, rb, true, false));
  RCHK(B.bind(c, b, rb, true, false));
  RCHK(Cc.bind(c, cc, rb, true, false));
  RCHK(R.bind(c, r, 2 * rb, true, false));
  RCHK(U.bind(c, u, 2 * rb, true, true));
  HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->s));
  hipLaunchKernelGGL(k_tridag_arrays, grid1(nh), dim3(256), 0, c->s, (const double *)A.dev, (const double *)B.dev,
                     (const double *)Cc.dev, (const cx<double> *)R.dev, (cx<double> *)U.dev, c->gam, g.lh, g.ny, n,
                     c->d_err);
  HIPCHK(hipGetLastError());
  RCHK(finish_all(c, {&A, &B, &Cc, &R, &U}));
  return check_err_flag(c);
}

// ------------------------------------------------------------------------------------
// frozen-coefficient steps and their discrete adjoint
// ------------------------------------------------------------------------------------
extern "C" int lp_set_frozen(lp_ctx *c, const void *nut, const void *ust) {
  if (!c) return set_err(LP_ERR_ARG, "lp_set_frozen: null context");
  if ((nut == nullptr) != (ust == nullptr))
    return set_err(LP_ERR_ARG, "lp_set_frozen: give both Nu_t and u*, or neither");
  if (nut && c->cfg.scal.s_flag)
    return set_err(LP_ERR_ARG, "lp_set_frozen: the frozen-coefficient step has no temperature (the wall law's "
                               "stability functions would not be frozen)");
  if (nut && (!is_device_ptr(nut) || !is_device_ptr(ust)))
    return set_err(LP_ERR_ARG, "lp_set_frozen: Nu_t and u* must be device arrays");
  c->frz_nut = nut;
  c->frz_ust = ust;
  return LP_OK;
}

namespace {
// the adjoint's binary64 scratch: NV volumes and NB 3/2-grid volumes, then lambda
constexpr int kAdjNV = 52, kAdjNB = 15;
}  // namespace

// main.f90:252-826 with the frozen Nu_t / u*, transposed (the oracle's or_step_adjoint,
// kernel for kernel): ub / rb hold on entry the adjoints of the step's u, v, w and of the
// RHS it leaves, on return those of the u, v, w it starts from and of the RHS_f it reads.
static int do_step_adjoint(lp_ctx *c, const double *nut, const double *ust, double *ub, double *rb, double *cb,
                           double *rcb) {
  const lp_grid &gg = c->g;
  const lp_config &cf = c->cfg;
  // the droplet gather's grid z covers (level, bin): at most 65 535, checked before any launch
  if ((long)gg.nz * cf.drop.npcon > 65535)
    return set_err(LP_ERR_ARG, "droplet adjoint: nz * npcon = %ld exceeds 65535", (long)gg.nz * cf.drop.npcon);
  const long nv = vol_el(c), np = plane_el(c);
  const long npb = (long)gg.ld_big * gg.ny2, nvb = npb * gg.nz;
  const long nh = (long)gg.lh * gg.ny;
  const size_t need = sizeof(double) * ((size_t)kAdjNV * nv + (size_t)kAdjNB * nvb) + sizeof(double) * 2 * nh * (gg.nz + 1);
  if (!c->adj_block) {
    RCHK(dalloc((void **)&c->adj_block, need));
    c->adj_bytes = need;
  }
  double *V[kAdjNV], *B[kAdjNB];
  for (int q = 0; q < kAdjNV; ++q) V[q] = c->adj_block + (size_t)q * nv;
  for (int q = 0; q < kAdjNB; ++q) B[q] = c->adj_block + (size_t)kAdjNV * nv + (size_t)q * nvb;
  cx<double> *lam = (cx<double> *)(c->adj_block + (size_t)kAdjNV * nv + (size_t)kAdjNB * nvb);
  AdjGeo ag{gg.nx, gg.ny, gg.nz, gg.ld, np};
  const dim3 gv = grid1(nv);
  const double dz = gg.dz, dt = cf.dt, cst = 1.0 / ((double)gg.nx * gg.ny);
  const int first = c->jt == 1 && !cf.initu;
  const int wall = cf.lbc_mom == LP_LBC_WALL;
  const double zl = 0.5 * dz - cf.d0, den = std::log(zl / cf.z0);
  const double *au = ub, *av = ub + nv, *aw = ub + 2 * nv;
  double *rbx = rb, *rby = rb + nv, *rbz = rb + 2 * nv;
  // ---- base intermediates from the context's u, v, w (left untouched) ----
  double *uf = V[0], *vf = V[1], *wf = V[2], *ux = V[3], *uy = V[4], *vx = V[5], *vy = V[6], *wx = V[7];
  double *wy = V[8], *uz = V[9], *vz = V[10], *wz = V[11], *om1 = V[12], *om2 = V[13], *om3 = V[14];
  HIPCHK(hipMemcpyAsync(uf, c->f[LP_U], sizeof(double) * nv, hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(vf, c->f[LP_V], sizeof(double) * nv, hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(wf, c->f[LP_W], sizeof(double) * nv, hipMemcpyDeviceToDevice, c->s));
  RCHK(do_filt_da(c, uf, ux, uy, gg.nz));
  RCHK(do_filt_da(c, vf, vx, vy, gg.nz));
  RCHK(do_filt_da(c, wf, wx, wy, gg.nz));
  RCHK(do_ddz_uv(c, uz, uf));
  RCHK(do_ddz_uv(c, vz, vf));
  RCHK(do_ddz_w(c, wz, wf));
  hipLaunchKernelGGL(k_adj_base, gv, dim3(256), 0, c->s, uf, vf, uz, vz, uy, vx, wx, wy, om1, om2, om3, ag,
                     1.0 / (zl * den), wall);
  HIPCHK(hipGetLastError());
  double *U1 = B[0], *U2 = B[1], *U3 = B[2], *O1 = B[3], *O2 = B[4], *O3 = B[5];
  RCHK(to_big(c, uf, gg.nz, U1));
  RCHK(to_big(c, vf, gg.nz, U2));
  RCHK(to_big(c, wf, gg.nz, U3));
  RCHK(to_big(c, om1, gg.nz, O1));
  RCHK(to_big(c, om2, gg.nz, O2));
  RCHK(to_big(c, om3, gg.nz, O3));
  // ---- projection, boundary values, RHS -= grad p ----
  double *usx = V[15], *usy = V[16], *usz = V[17], *apx = V[18], *apy = V[19], *apz = V[20];
  double *arx = V[21], *ary = V[22], *arz = V[23], *aft = V[24], *aftv = V[25], *ap = V[26];
  hipLaunchKernelGGL(k_adj_project, gv, dim3(256), 0, c->s, au, av, aw, rbx, rby, rbz, usx, usy, usz, apx, apy, apz,
                     arx, ary, arz, aft, aftv, ag, dt, cf.use_mean_p_force, cf.lbc_mom == LP_LBC_STRESS_FREE ? 1 : 0);
  hipLaunchKernelGGL(k_adj_dpdz, gv, dim3(256), 0, c->s, apz, ap, ag, 1.0 / dz);
  HIPCHK(hipGetLastError());
  // ---- pressure: c2r^T of p, dpdx, dpdy, the transposed solves, r2c^T of H, wall, lid ----
  const long ps = c->sp * gg.ny;
  RCHK(r2c_small(c, ap, gg.nz, 1.0, c->spec[0]));
  RCHK(r2c_small(c, apx, gg.nz - 1, 1.0, c->spec[1]));
  RCHK(r2c_small(c, apy, gg.nz - 1, 1.0, c->spec[2]));
  for (int q = 0; q < 3; ++q) {
    ColOut o{c->spec[q], M_ONE};
    RCHK(col_small(c, c->spec[q], q == 0 ? gg.nz : gg.nz - 1, 1, 0, 1, &o));
  }
  const dim3 gm((unsigned)((nh + 63) / 64));
  hipLaunchKernelGGL(k_press_adj, gm, dim3(64), 0, c->s, (const cx<double> *)c->spec[0],
                     (const cx<double> *)c->spec[1], (const cx<double> *)c->spec[2], c->kx, c->ky,
                     (cx<double> *)c->Hs[0], (cx<double> *)c->Hs[1], (cx<double> *)c->Hs[2], (cx<double> *)c->bw,
                     (cx<double> *)c->tw, c->gam, lam, gg.lh, gg.ny, gg.nz, c->sp, dz, cst / (1.5 * dt), cst);
  HIPCHK(hipGetLastError());
  double *hx = V[27], *hy = V[28], *hz = V[29], *bwb = V[30], *twb = V[31];
  double *hout[3] = {hx, hy, hz};
  for (int q = 0; q < 3; ++q) {
    ColOut o{c->Hs[q], M_ONE};
    RCHK(col_small(c, c->Hs[q], gg.nz - 1, 0, 1, 1, &o));
    RCHK(c2r_small(c, c->Hs[q], gg.lh, gg.nz - 1, hout[q], F64));
  }
  void *wl[2] = {c->bw, c->tw};
  double *wlo[2] = {bwb, twb};
  for (int q = 0; q < 2; ++q) {
    ColOut o{wl[q], M_ONE};
    RCHK(col_small(c, wl[q], 1, 0, 1, 1, &o));
    RCHK(c2r_small(c, wl[q], gg.lh, 1, wlo[q], F64));
  }
  (void)ps;
  // ---- AB2, RHS assembly, Euler start ----
  double *afu = V[32], *afv = V[33], *afw = V[34], *acx = V[35], *acy = V[36], *acz = V[37];
  double *adx = V[38], *ady = V[39], *adz = V[40];
  hipLaunchKernelGGL(k_adj_ab2, gv, dim3(256), 0, c->s, usx, usy, usz, hx, hy, hz, arx, ary, arz, aft, aftv, bwb, twb,
                     afu, afv, afw, acx, acy, acz, adx, ady, adz, rbx, rby, rbz, ag, dt, first);
  HIPCHK(hipGetLastError());
  // ---- droplets (main.f90:360-372): update, buoyancy and transport transposed, every
  // bin in each launch; the face-velocity adjoints reach u, v, w after the strain rates ----
  const lp_droplet_config &D = cf.drop;
  const bool drops = D.npcon > 0 && c->jt_total >= D.pcon_init;
  double *dmu = nullptr, *dac = nullptr, *awi = nullptr, *aui = nullptr, *avi = nullptr;
  DropGeo dg{};
  if (drops) {
    dg = drop_geo(c);
    const size_t fz = (size_t)dg.fp * gg.nz, nb = (size_t)D.npcon;
    // mu, ac, awi (a volume per bin) and aui, avi (a face array per bin)
    if (!c->adj_drop) RCHK(dalloc((void **)&c->adj_drop, sizeof(double) * nb * (3 * (size_t)nv + 2 * fz)));
    dmu = c->adj_drop; dac = dmu + nb * nv; awi = dac + nb * nv; aui = awi + nb * nv; avi = aui + nb * fz;
    // base face velocities from the filtered u, v and dudx, dvdy; Kc_t from the frozen Nu_t
    hipLaunchKernelGGL(k_face_u<double>, dim3((unsigned)(((long)gg.ny * gg.nz + 3) / 4)), dim3(256), 0, c->s, uf, ux,
                       c->ui0, dg, gg.dx);
    hipLaunchKernelGGL(k_face_v<double>, dim3((unsigned)((gg.nx + 63) / 64), (unsigned)gg.nz), dim3(64 * kFaceSeg), 0,
                       c->s, vf, vy, c->vi0, dg, gg.dy);
    hipLaunchKernelGGL(k_kc<double>, grid1(dg.nv), dim3(256), 0, c->s, nut, (double *)c->f[LP_MAGS], c->kcell, dg,
                       D.Sc);
    const long nfill = (long)(nb * (nv + 2 * fz));
    hipLaunchKernelGGL(k_fill<double>, grid1(nfill), dim3(256), 0, c->s, awi, nfill, 0.0);
    HIPCHK(hipGetLastError());
    const int euler = (c->jt_total == D.pcon_init || c->jt_total == D.src_jt0) && !D.init_pcon;
    DropAdjP P{1.0 / gg.dx, 1.0 / gg.dy, 1.0 / gg.dz, 1.0 / (gg.dx * gg.dx), 1.0 / (gg.dy * gg.dy),
               1.0 / (gg.dz * gg.dz), D.bc_x, D.bc_y, D.settling};
    const double *PC = (const double *)c->f[LP_PCON];
    hipLaunchKernelGGL(k_adj_drop_update, grid1(dg.nv * D.npcon), dim3(256), 0, c->s, (const double *)cb, rcb,
                       (const double *)arz, (const double *)c->d_vfac, D.active, dmu, dac, dg, dt, euler);
    // 64 x 4 cells per block, grid z over (level, bin) (nz * npcon checked on entry)
    hipLaunchKernelGGL(k_adj_drop_xyz, dim3((unsigned)((gg.nx + 63) / 64), (unsigned)((gg.ny + 3) / 4),
                       (unsigned)(gg.nz * D.npcon)), dim3(64, 4), 0, c->s, PC, (const double *)dmu, dac,
                       (const double *)c->ui0, (const double *)c->vi0, (const double *)wf, aui, avi, awi,
                       (const double *)c->kcell, (const double *)c->d_ws, dg, P);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cb, dac, sizeof(double) * nb * nv, hipMemcpyDeviceToDevice, c->s));
  }
  // ---- convection: P^T, the products transposed, T^T ----
  double *pbx = B[6], *pby = B[7], *pbz = B[8];
  RCHK(to_big(c, acx, gg.nz, pbx));
  RCHK(to_big(c, acy, gg.nz, pby));
  RCHK(to_big(c, acz, gg.nz, pbz));
  double *aU1 = B[9], *aU2 = B[10], *aU3 = B[11], *aO1 = B[12], *aO2 = B[13], *aO3 = B[14];
  const double tt = ((double)gg.nx * gg.ny) / ((double)gg.nx2 * gg.ny2);
  hipLaunchKernelGGL(k_adj_products, grid1(nvb), dim3(256), 0, c->s, pbx, pby, pbz, U1, U2, U3, O1, O2, O3, aU1, aU2,
                     aU3, aO1, aO2, aO3, npb, gg.nz, tt);
  HIPCHK(hipGetLastError());
  double *t1 = V[41], *ao1 = V[42], *ao2 = V[43], *ao3 = V[44];
  double *aUb[3] = {aU1, aU2, aU3}, *afq[3] = {afu, afv, afw};
  for (int q = 0; q < 3; ++q) {  // P^T = (nx2 ny2)/(nx ny) T: truncation scaled by 1/(nx ny)
    RCHK(from_big(c, aUb[q], gg.nz, cst, t1));
    hipLaunchKernelGGL(k_add_plane, gv, dim3(256), 0, c->s, afq[q], t1, nv);
  }
  RCHK(from_big(c, aO1, gg.nz, cst, ao1));
  RCHK(from_big(c, aO2, gg.nz, cst, ao2));
  RCHK(from_big(c, aO3, gg.nz, cst, ao3));
  // ---- stress divergences: the x, y derivatives of their adjoints ----
  double *dxa = V[45], *dya = V[46], *dxb = V[47], *dyb = V[48], *dxc = V[49], *dyc = V[50];
  const int mxy[2] = {M_IKX, M_IKY};
  void *oa[2] = {dxa, dya}, *ob[2] = {dxb, dyb}, *oc[2] = {dxc, dyc};
  RCHK(spectral_apply(c, adx, gg.nz, cst, 2, mxy, oa));
  RCHK(spectral_apply(c, ady, gg.nz, cst, 2, mxy, ob));
  RCHK(spectral_apply(c, adz, gg.nz, cst, 2, mxy, oc));
  // ---- stresses with the frozen Nu_t -> strain-rate adjoints (reuse the base volumes) ----
  double *aS11 = V[15], *aS12 = V[16], *aS13 = V[17], *aS22 = V[18], *aS23 = V[19], *aS33 = V[20];
  double *wtx = V[21], *wty = V[22];
  hipLaunchKernelGGL(k_adj_tau, gv, dim3(256), 0, c->s, adx, ady, adz, dxa, dya, dxb, dyb, dxc, dyc, nut, ust, aS11,
                     aS12, aS13, aS22, aS23, aS33, wtx, wty, ag, 1.0 / dz, cf.vonk / den, wall);
  HIPCHK(hipGetLastError());
  if (wall) {  // the wall law's test filter (symmetric) on the txz(1), tyz(1) adjoints
    const int mg[1] = {M_G1};
    void *o1[1] = {t1};
    RCHK(spectral_apply(c, wtx, 1, 1.0, 1, mg, o1));
    hipLaunchKernelGGL(k_add_plane, grid1(np), dim3(256), 0, c->s, afu, t1, np);
    RCHK(spectral_apply(c, wty, 1, 1.0, 1, mg, o1));
    hipLaunchKernelGGL(k_add_plane, grid1(np), dim3(256), 0, c->s, afv, t1, np);
    HIPCHK(hipGetLastError());
  }
  double *aux = V[3], *auy = V[4], *avx = V[5], *avy = V[6], *awx = V[7], *awy = V[8];
  hipLaunchKernelGGL(k_adj_strain, gv, dim3(256), 0, c->s, aS11, aS12, aS13, aS22, aS23, aS33, ao1, ao2, ao3, aux, auy,
                     avx, avy, awx, awy, afu, afv, afw, ag, 1.0 / dz, 1.0 / (zl * den), wall);
  HIPCHK(hipGetLastError());
  if (drops) {  // the droplet faces read the filtered u, v, w and dudx, dvdy
    hipLaunchKernelGGL(k_adj_face_w, gv, dim3(256), 0, c->s, (const double *)awi, afw, dg);
    // the bin sums of the u and v face adjoints go where mu, ac and awi were (dead now:
    // 3 npcon nv >= 2 fp nz)
    const long fz = (long)dg.fp * gg.nz;
    double *lu = dmu, *lv = dmu + fz;
    hipLaunchKernelGGL(k_adj_face_bins, grid1(fz), dim3(256), 0, c->s, (const double *)aui, (const double *)avi, lu, lv,
                       dg);
    hipLaunchKernelGGL(k_adj_face_line<false>, dim3((unsigned)(((long)gg.ny * gg.nz + 3) / 4)), dim3(256), 0, c->s,
                       (const double *)lu, aux, afu, dg, gg.dx);
    hipLaunchKernelGGL(k_adj_face_line<true>, dim3((unsigned)(((long)gg.nx * gg.nz + 3) / 4)), dim3(256), 0, c->s,
                       (const double *)lv, avy, afv, dg, gg.dy);
    HIPCHK(hipGetLastError());
  }
  // ---- filt_da transposed: F(af) - Dx(a*x) - Dy(a*y) ----
  double *ff = V[9], *gx = V[10], *gy = V[11];
  double *axq[3] = {aux, avx, awx}, *ayq[3] = {auy, avy, awy};
  const int m1[1] = {M_ONE}, mx[1] = {M_IKX}, my[1] = {M_IKY};
  for (int q = 0; q < 3; ++q) {
    void *of[1] = {ff}, *ox[1] = {gx}, *oy[1] = {gy};
    RCHK(spectral_apply(c, afq[q], gg.nz, cst, 1, m1, of));
    RCHK(spectral_apply(c, axq[q], gg.nz, cst, 1, mx, ox));
    RCHK(spectral_apply(c, ayq[q], gg.nz, cst, 1, my, oy));
    hipLaunchKernelGGL(k_adj_combine, gv, dim3(256), 0, c->s, ub + (size_t)q * nv, ff, gx, gy, ag);
    HIPCHK(hipGetLastError());
  }
  return LP_OK;
}

extern "C" int lp_step_adjoint2(lp_ctx *c, const double *nut, const double *ust, double *ub, double *rb, double *cb,
                                double *rcb) {
  if (!c || !nut || !ust || !ub || !rb) return set_err(LP_ERR_ARG, "lp_step_adjoint: null argument");
  if (c->fdt != F64)
    return set_err(LP_ERR_ARG, "lp_step_adjoint runs in binary64: use an LP_FP64 context (the base states of a "
                               "binary32 run carry over, u, v, w being binary64 in both)");
  if (c->cfg.scal.s_flag)
    return set_err(LP_ERR_ARG, "lp_step_adjoint: the temperature (theta and its buoyancy) is not differentiated");
  const lp_droplet_config &D = c->cfg.drop;
  if (D.npcon > 0) {
    if (!cb || !rcb) return set_err(LP_ERR_ARG, "lp_step_adjoint: droplets need the PCon and RHS_PCon adjoints");
    if (D.inertia) return set_err(LP_ERR_ARG, "lp_step_adjoint: droplet inertia (dudt in the face velocity) is not differentiated");
    if (D.pbe && D.pbe_feedback) return set_err(LP_ERR_ARG, "lp_step_adjoint: the breakup feedback is not differentiated");
  }
  HIPCHK(hipSetDevice(c->dev));
  if (!is_device_ptr(nut) || !is_device_ptr(ust) || !is_device_ptr(ub) || !is_device_ptr(rb) ||
      (D.npcon > 0 && (!is_device_ptr(cb) || !is_device_ptr(rcb))))
    return set_err(LP_ERR_ARG, "lp_step_adjoint: device arrays expected");
  RCHK(do_step_adjoint(c, nut, ust, ub, rb, cb, rcb));
  HIPCHK(hipStreamSynchronize(c->s));
  return LP_OK;
}

extern "C" int lp_step_adjoint(lp_ctx *c, const double *nut, const double *ust, double *ub, double *rb) {
  return lp_step_adjoint2(c, nut, ust, ub, rb, nullptr, nullptr);
}
