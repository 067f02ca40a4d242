// lp_internal.hpp -- host-side interface between the C ABI layer (lp_api.hip) and the
// kernel launchers (lp_fft.hip, lp_physics.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lp {

enum Dt { F32 = 0, F64 = 1 };

// twiddle tables exp(-2 pi i j / N), j < N, in both precisions
struct Twiddle {
  int n;
  const void *d;  // double2[n]
  const void *f;  // float2[n]
};
const void *tw_ptr(int n, Dt dt);  // device pointer; aborts the launch (nullptr) if missing

// multiplier applied per spectral output of a column operation
enum Mult { M_ONE = 0, M_IKX = 1, M_IKY = 2, M_G1 = 3, M_G2 = 4 };

struct SpecTables {
  const double *kx, *ky, *G1, *G2;  // (lh, ny) of the small grid
  int lh, ny;                        // small-grid sizes
};

// Input of a row pass: din is the type of the rows (-1: the transform's type; F64 into an
// F32 transform reads binary64 velocities).  shift (F64 into F32 only): each plane enters
// less its first value, stored in shift_out[plane] when shift_out is given; only the
// plane's kx = ky = 0 mode moves, by that value.
struct RowIn {
  int din = -1;
  bool shift = false;
  double *shift_out = nullptr;
};

// Row pass: real rows of length 2M (scaled by `scale`) -> complex half spectra
// (M+1 values, only the first ncols_out written).
hipError_t row_r2c(Dt dt, int M, const void *in, long in_pitch, long in_plane,
                   int rows, int nplanes, double scale, void *out, long out_pitch,
                   long out_plane, int ncols_out, hipStream_t s, RowIn ri = {});

// Row pass: complex half spectra (ncols_in columns present, rest zero) -> real rows
// of length 2M; columns [2M, out_pitch) set to zero.  Compute type dtc, store type dto.
// ikx (optional, M+1 values, zero at kx = 0 and kx = M): the spectra are multiplied by
// i*ikx[kx] first, so the rows are the x-derivative of what they would be without it.
// shift (optional, dtc F32 -> dto F64): shift[plane] is added to every value of the plane
// in binary64, undoing row_r2c's shift.
hipError_t row_c2r(Dt dtc, Dt dto, int M, const void *in, long in_pitch, long in_plane,
                   int ncols_in, int rows, int nplanes, void *out, long out_pitch,
                   long out_plane, hipStream_t s, const double *ikx = nullptr,
                   const double *shift = nullptr);

// row_r2c of a vorticity component formed from two derivative volumes (comp 0: w_y - v_z
// from a = dwdy, b = dvdz; 1: u_z - w_x from a = dwdx, b = dudz; 2: v_x - u_y from
// a = dvdx, b = dudy), the x-derivatives of w averaged onto the uvp node on plane 0.
hipError_t row_r2c_vort(Dt dt, int M, int comp, const void *a, const void *b, long in_pitch, long in_plane,
                        int rows, int nplanes, double scale, void *out, long out_pitch, long out_plane,
                        int ncols_out, hipStream_t s);

struct ColOut {
  void *ptr;
  int mult;
  long pitch = 0;  // overrides of the call's out_pitch / out_plane / ncols (0: keep)
  long plane = 0;
  int ncols = 0;
  double scale = 1.0;  // padmode 1 and 2 only: factor on the output spectrum
};

// Column pass over ncols kx-columns of nplanes spectra:
//  fwd: forward complex DFT of length nyi along y;
//  map: padmode 0 (same grid, Nyquist row/column killed, multiplier applied),
//       1 (zero-pad ny -> ny2 as padd.f90), 2 (truncate ny2 -> ny as unpadd.f90);
//  inv: inverse DFT of length nyo.
// Timing events taken by the dispatch of a launch itself (hipExtLaunchKernel): start at
// the first wave, stop at the end of the kernel, with no marker packet in the stream.
// Either may be null.
struct LaunchEv {
  hipEvent_t start = nullptr, stop = nullptr;
};

hipError_t col_op(Dt dt, int nyi, int nyo, int padmode, int fwd, int inv, const void *in,
                  long in_pitch, long in_plane, int ncols, int nplanes, int nout,
                  const ColOut *outs, long out_pitch, long out_plane, const SpecTables &tab,
                  hipStream_t s, LaunchEv ev = {});

// col_op (same grid, forward + inverse) on the columns col0..ncols-1 only, one column per
// block: the few columns past the last full group of a block-wide launch.
hipError_t col_op_tail(Dt dt, int ny, const void *in, long in_pitch, long in_plane, int col0, int ncols,
                       int nplanes, int nout, const ColOut *outs, long out_pitch, long out_plane,
                       const SpecTables &tab, hipStream_t s, LaunchEv ev = {});

// Column pass of the divergence terms: out = inverse_y( i kx fwd_y(A) + i ky fwd_y(B) )
// with the Nyquist row and column zeroed; A, B, out share pitch / plane; ny == tab.ny.
hipError_t col_div(Dt dt, int ny, const void *A, const void *B, long pitch, long plane, int ncols, int nplanes,
                   void *out, const SpecTables &tab, hipStream_t s);

// row_c2r of small spectra followed by the z-derivative term of divstress: is_w = 0
// (divstress_uv) adds cst (tz(k+1) - tz(k)) and writes BOGUS on the top level; is_w = 1
// (divstress_w) adds cst (tz(k) - tz(k-1)) on levels 1..nplanes-2.  Compute and store
// type dt; tz has the layout of out.
hipError_t row_c2r_div(Dt dt, int M, const void *in, long in_pitch, long in_plane, int ncols_in, int rows,
                       int nplanes, void *out, long out_pitch, long out_plane, const void *tz, double cst,
                       int is_w, hipStream_t s);

bool fft_len_supported(int n);

// record the message that lp_last_error() returns; returns code
int set_error(int code, const char *fmt, ...);

}  // namespace lp
