// lp_lag.hpp -- launchers of the fused Lagrangian-model passes (lp_lag.hip):
//   lag_r2c       : the 21 test-filter inputs of lagrange_Sdep.f90:129-155,192-242,
//                   formed on the fly from u, v, w, S_ij, |S| and transformed along x
//                   (only the leading kx columns the filters keep are written)
//   lag_c2r_contract : x back-transforms of the filtered fields in LDS and the Germano
//                   contractions (lagrange_Sdep.f90:257-326), two values per point and
//                   width, stored in the field type they are computed in
//   lag_relax     : per point, the backtrace interpolation (interpolag_Sdep.f90:62-365)
//                   and the relaxation (lagrange_Sdep.f90:327-434)
// Between lag_r2c and lag_c2r_contract: col_op over all 21 fields at once.
#pragma once
#include "lp_internal.hpp"
#include "lp_kernels.hpp"

namespace lp {

struct LagInputs {
  const void *u, *v, *w;  // binary64 in both precisions
  const void *S[6];
  const void *magS;  // |S| of the strain rates (same value sgs_stag computes)
};

struct LagInterp {
  double cs;          // cs_count
  double sx, sy, sz;  // dt cs / dx, dt cs / dy, dt cs / dz
};

// spec: [21][nz][ny][pitch] complex of the field type, the leading nc columns written
hipError_t lag_r2c(Dt dt, const LagInputs &in, Geo g, void *spec, int nc, long pitch, hipStream_t s);

// bar/hat: y-filtered spectra [21][nz][ny][p1|p2] (n1|n2 columns used); C: L.M, M.M,
// Q.N, N.N (volumes of the field type: the contractions are formed in it, so storing
// them in it loses nothing)
hipError_t lag_c2r_contract(Dt dt, const void *bar, const void *hat, int n1, int n2, long p1, long p2,
                            void *const C[4], Geo g, const LagParams &P, hipStream_t s);

// C as above; Fold are read (gathered), Fnew written.  ul, vl, wl: the lag velocities
// (vel_f64 false: the field type) or u, v, w themselves (vel_f64 true: binary64).
hipError_t lag_relax(Dt dt, void *const C[4], const double *const Fold[4], double *const Fnew[4], const void *ul,
                     const void *vl, const void *wl, void *Cs, void *Beta, Geo g, const LagParams &P,
                     const LagInterp &I, hipStream_t s, bool vel_f64);

}  // namespace lp
