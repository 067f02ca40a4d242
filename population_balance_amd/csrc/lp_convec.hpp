// lp_convec.hpp -- launcher of the fused 3/2-grid row pass of convec.f90 (lp_convec.hip):
// x back-transforms of the six padded fields, the cross products
// (convec.f90:138-158, 180-221, 241-279) and the x forward transforms of the three
// products, one (y, level-chunk) of the 3/2 grid per block.  The padded fields and the
// products exist only as half spectra along x (kx < nx/2), so no 3/2-grid physical
// volume is ever stored.
#pragma once
#include "lp_internal.hpp"

namespace lp {

// B: y-padded spectra [6][nz][ny2][pitch] of u1 u2 u3 w1 w2 w3 (kx < nx/2 present);
// C: product spectra [3][nz][ny2][pitch] (cx, cy, cz; levels 0..nz-2 written, kx < nx/2).
// M = nx2/2; cb = the dealiasing normalisation folded into the products.
hipError_t convec_rows(Dt dt, int M, const void *B, void *C, long pitch, int ny2, int nz, int ncols, double cb,
                       hipStream_t s);

}  // namespace lp
