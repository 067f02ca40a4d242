// lp_lag.hpp -- launchers of the fused Lagrangian-model passes (lp_lag.hip):
//   lag_r2c       : the 21 test-filter inputs of lagrange_Sdep.f90:129-155,192-242,
//                   formed on the fly from u, v, w, S_ij, |S| and transformed along x
//                   (only the leading kx columns the filters keep are written)
//   lag_relax     : per point, the backtrace interpolation (interpolag_Sdep.f90:62-365),
//                   the Germano contractions and the relaxation (lagrange_Sdep.f90:257-434)
//                   from the 42 filtered volumes
// Between them: col_op over all 21 fields at once, then row_c2r into the 42 volumes.
#pragma once
#include "lp_internal.hpp"
#include "lp_kernels.hpp"

namespace lp {

struct LagInputs {
  const void *u, *v, *w;
  const void *S[6];
  const void *magS;  // |S| of the strain rates (same value sgs_stag computes)
};

struct LagInterp {
  double cs, dt, dx, dy, dz;
};

// spec: [21][nz][ny][pitch] complex of the field type, the leading nc columns written
hipError_t lag_r2c(Dt dt, const LagInputs &in, Geo g, void *spec, int nc, long pitch, hipStream_t s);

// bar, hat: the 21 filtered fields as volumes (field q at q*ld*ny*nz); Fold are read
// (gathered), Fnew written
hipError_t lag_relax(Dt dt, const void *bar, const void *hat, const double *const Fold[4], double *const Fnew[4],
                     const void *ul, const void *vl, const void *wl, void *Cs, void *Beta, Geo g,
                     const LagParams &P, const LagInterp &I, hipStream_t s);

}  // namespace lp
