// lp_fft.hip -- batched plane FFT passes for gfx950.
//
// A 2-D real transform of an (ld,ny) plane (FFTW2 rfftwnd in-place layout,
// fft.f90:35-46) is split into
//   row_r2c : each x-row of 2M reals -> M+1 complex (half-length complex FFT + split)
//   col_op  : per kx column, complex FFT along y, a spectral map, inverse FFT along y
//   row_c2r : M+1 complex -> 2M reals (merge + half-length inverse FFT), using only the
//             real parts of the kx = 0 and kx = M terms (FFTW2 halfcomplex semantics)
// Every transform is a register-resident four-step FFT (lp_fft4.hpp) with one LDS
// exchange; loads and stores go straight between HBM and registers.
#include <hip/hip_ext.h>

#include "lp_fft4.hpp"
#include "lp_internal.hpp"

namespace lp {

// c2r block: 256 threads; 128 in binary64, whose slots (16-byte values) would otherwise
// take 88 KB of LDS per block and leave one block per CU
template <typename T> __host__ __device__ constexpr int row_block() { return sizeof(T) == 8 ? 128 : 256; }

// XCD-aware plane order for the binary64 passes.  Workgroup ids are dealt round-robin to
// the 8 XCDs, each with its own L2.  The column pass and the back-transform give plane p to
// XCD p mod 8, so that a plane written by one is read by the other on the same XCD; the
// column pass takes an XCD's planes last first, the back-transform first first (the
// column pass wrote those last).  In binary32 the same order measured slightly slower
// (its planes are half the size, and more of them are served from the cache behind L2).
// xcd_planes(n, j): the planes of XCD j.
__host__ __device__ constexpr int xcd_planes(int nplanes, int j) { return (nplanes - j + 7) / 8; }
template <typename T> __host__ __device__ constexpr bool xcd_order() { return sizeof(T) == 8; }
static constexpr int kR2cBlock = 128;  // r2c: persistent blocks of 128 threads

template <typename T> __host__ __device__ constexpr bool r2c_persistent() { return sizeof(T) == 4; }

// grid of the r2c pass: persistent (8 blocks per compute unit, at most one per group) or
// one block per group
static unsigned r2c_grid(long ngroups, bool persistent) {
  if (!persistent) return (unsigned)ngroups;
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      cus = n;
    else
      cus = 256;
  }
  const long g = 8L * cus;
  return (unsigned)(ngroups < g ? ngroups : g);
}

// ---------------------------------------------------------------------------------
// row_r2c: real rows of 2M values -> M+1 complex (half-length complex FFT + split)
// ---------------------------------------------------------------------------------
// PRO selects the input: 0 the rows of `in`; 1..3 a vorticity component formed from two
// derivative volumes a, b (convec.f90:100-133): 1 w_y - v_z (a = dwdy, b = dvdz), 2
// u_z - w_x (a = dwdx, b = dudz), 3 v_x - u_y (a = dvdx, b = dudy); on plane 0 (jz = 1)
// w_y and w_x are averaged with plane 1 onto the uvp node.  a, b have in's layout.
// TI is the type of `in` (binary64 velocities into a binary32 transform).  SH: every row
// of a plane enters shifted by the plane's first value, (in - in[plane, 0, 0]) * scale,
// which only moves the plane's kx = ky = 0 mode; the row y = 0 stores that value in
// shift_out[plane] when shift_out is given.
template <typename T, int M, int PRO = 0, typename TI = T, bool SH = false>
__global__ __launch_bounds__(kR2cBlock) void k_row_r2c(const TI *__restrict__ in, long in_pitch, long in_plane,
                                                       int rows, int nplanes, TI scale, cx<T> *__restrict__ out,
                                                       long out_pitch, long out_plane, int ncols_out,
                                                       const cx<T> *__restrict__ twM,
                                                       const cx<T> *__restrict__ twN,
                                                       const T *__restrict__ pa = nullptr,
                                                       const T *__restrict__ pb = nullptr,
                                                       double *__restrict__ shift_out = nullptr) {
  constexpr int P = row_lanes(M);
  using F = Fft4<T, M, P, -1>;
  constexpr int Q = F::Q;
  constexpr int RB = kR2cBlock / P;
  constexpr int SS = slot_stride(F::SCR > M ? F::SCR : M);
  __shared__ cx<T> lds[RB * SS];
  __shared__ cx<T> tws[2 * M + 1];  // twM[0..M), twN[0..M]: read from LDS by every group
  const int t = threadIdx.x % P, r = threadIdx.x / P;
  const long total = (long)rows * nplanes;
  const long ngroups = (total + RB - 1) / RB;
  cx<T> *sc = lds + r * SS;
  for (int i = threadIdx.x; i <= 2 * M; i += kR2cBlock) tws[i] = i < M ? twM[i] : twN[i - M];
  __syncthreads();
  using v2 = typename V2<T>::t;
  // input of row group g (rows beyond the last read row 0 and are not stored)
  auto load = [&](long g, cx<T> (&v)[Q]) {
    const long grow = g * RB + r;
    const bool valid = grow < total;
    const long plane = valid ? grow / rows : 0, y = valid ? grow % rows : 0;
    if constexpr (PRO == 0) {
      using vi2 = typename V2<TI>::t;
      const TI *src = in + plane * in_plane + y * in_pitch;
      TI s0 = TI(0);
      if constexpr (SH) s0 = in[plane * in_plane];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const vi2 p = *reinterpret_cast<const vi2 *>(src + 2 * (t + P * q));
        if constexpr (SH)
          v[q] = mk<T>((T)((p.x - s0) * scale), (T)((p.y - s0) * scale));
        else
          v[q] = mk<T>((T)(p.x * scale), (T)(p.y * scale));
      }
      if constexpr (SH) {
        if (shift_out && valid && y == 0 && t == 0) shift_out[plane] = (double)s0;
      }
    } else {
      const long o = plane * in_plane + y * in_pitch;
      const long up = (PRO != 3 && plane == 0) ? in_plane : 0;  // clamped: same row above jz = 1
      const T h = T(0.5);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const long x = o + 2 * (t + P * q);
        const v2 a0 = *reinterpret_cast<const v2 *>(pa + x);
        const v2 b0 = *reinterpret_cast<const v2 *>(pb + x);
        v2 rr;
        if constexpr (PRO == 3) {
          rr = v2{a0.x - b0.x, a0.y - b0.y};
        } else {
          const v2 a1 = *reinterpret_cast<const v2 *>(pa + x + up);
          const v2 av{h * (a0.x + a1.x), h * (a0.y + a1.y)};
          const v2 aa = plane == 0 ? av : a0;
          rr = PRO == 1 ? v2{aa.x - b0.x, aa.y - b0.y} : v2{b0.x - aa.x, b0.y - aa.y};
        }
        v[q] = mk<T>(rr.x * scale, rr.y * scale);
      }
    }
  };
  // binary32: persistent blocks, the next group's loads issued before this group's
  // transform; binary64 (twice the registers per value): one group per block
  constexpr bool PF = r2c_persistent<T>();
  long g = blockIdx.x;
  cx<T> v[Q];
  if (PF && g < ngroups) load(g, v);
  for (; g < ngroups; g += gridDim.x) {
    cx<T> vn[Q];
    if (!PF) load(g, v);
    else if (g + gridDim.x < ngroups) load(g + gridDim.x, vn);
    const long grow = g * RB + r;
    const bool valid = grow < total;
    const long plane = valid ? grow / rows : 0, y = valid ? grow % rows : 0;
    F::stage1(v, t, sc, tws);
    slot_sync<P>();
    cx<T> a[F::G][P];
    bool ok[F::G];
#pragma unroll
    for (int j = 0; j < F::G; ++j) ok[j] = F::stage2(j, t, sc, a[j]);
    slot_sync<P>();
#pragma unroll
    for (int j = 0; j < F::G; ++j)
      if (ok[j]) {
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) sc[t + P * j + Q * k1] = a[j][k1];
      }
    slot_sync<P>();
    // split: X[k] = (Z[k] + conj Z[M-k])/2 - i/2 w^k (Z[k] - conj Z[M-k]), w = exp(-2 pi i/2M)
    if (valid) {
      cx<T> *dst = out + plane * out_plane + y * out_pitch;
      for (int k = t; k <= M && k < ncols_out; k += P) {
        cx<T> za = sc[k % M], zb = conjg(sc[(M - k) % M]);
        cx<T> e = za + zb, o = za - zb;
        cx<T> wo = cmul(tws[M + k], o);
        dst[k] = mk<T>(T(0.5) * (e.x + wo.y), T(0.5) * (e.y - wo.x));
      }
    }
    if (!PF) break;
    slot_sync<P>();
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = vn[q];
  }
}

// merge of the half spectrum into the M-point complex input of a 2M-point real inverse
// transform: Y[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k]), w = exp(-2 pi i/2M)
template <typename T> __device__ __forceinline__ cx<T> merge_c2r(cx<T> xa, cx<T> xm, cx<T> w) {
  const cx<T> xb = conjg(xm);
  const cx<T> e = xa + xb, o = xa - xb;
  const cx<T> wo = cmul(conjg(w), o);
  return mk<T>(e.x - wo.y, e.y + wo.x);
}

// ---------------------------------------------------------------------------------
// row_c2r: M+1 complex (first ncols_in present; imaginary parts of kx = 0 and kx = M
// ignored, as FFTW2's halfcomplex c2r) -> 2M reals, columns [2M, out_pitch) zeroed
// ---------------------------------------------------------------------------------
// EPI selects what is stored: 0 the transform; 1 (divstress_uv.f90:88-89,116-125) the
// transform plus dz^-1 (tz(k+1) - tz(k)), BOGUS on the top level; 2 (divstress_w.f90:35,
// 46-81 with the wall at level 1 and no lid term) the transform plus dz^-1 (tz(k) - tz(k-1))
// on levels 1..nz-2.  The padding columns are zero (BOGUS on the uv top level).  tz has the
// layout of out.  ikx (optional): the input is multiplied by i*ikx[kx] first, which turns
// the (kx, y) spectrum of f into that of df/dx (ddx.f90:15-24; kx is uniform along y).
// shift (optional, EPI 0): shift[plane] is added to every stored value of the plane in
// binary64 (the inverse of k_row_r2c's SH; padding columns stay zero).
template <typename T, typename TO, int M, int EPI, bool IKX = false>
__global__ __launch_bounds__(row_block<T>()) void k_row_c2r(const cx<T> *__restrict__ in, long in_pitch, long in_plane,
                                                       int ncols_in, int rows, int nplanes,
                                                       TO *__restrict__ out, long out_pitch, long out_plane,
                                                       const cx<T> *__restrict__ twM,
                                                       const cx<T> *__restrict__ twN,
                                                       const TO *__restrict__ tz, TO cst,
                                                       const double *__restrict__ ikx,
                                                       const double *__restrict__ shift, int xorder) {
  constexpr int P = row_lanes(M);
  using F = Fft4<T, M, P, +1>;
  constexpr int Q = F::Q;
  constexpr int RB = row_block<T>() / P;
  constexpr int SS = slot_stride(F::SCR);
  __shared__ cx<T> lds[RB * SS];
  __shared__ cx<T> tws[2 * M];  // twM[0..M), twN[0..M): read from LDS by every row
  const int t = threadIdx.x % P, r = threadIdx.x / P;
  long blk = blockIdx.x;
  if (xorder) {  // XCD order, first-written planes first; the grid has 8 max_j(planes of j) gpp blocks
    const int gpp = rows / RB, xj = (int)(blockIdx.x & 7u);
    const long xi = blockIdx.x >> 3, kq = xi / gpp;
    if (kq >= xcd_planes(nplanes, xj)) return;  // uniform over the block
    blk = (long)(xj + 8 * kq) * gpp + xi % gpp;
  }
  const long grow = blk * RB + r;
  const bool valid = grow < (long)rows * nplanes;
  const long plane = valid ? grow / rows : 0, y = valid ? grow % rows : 0;
  cx<T> *sc = lds + r * SS;
  const cx<T> *src = in + plane * in_plane + y * in_pitch;
  for (int i = threadIdx.x; i < 2 * M; i += row_block<T>()) tws[i] = i < M ? twM[i] : twN[i - M];
  __shared__ T ikxs[IKX ? M + 1 : 1];  // the multipliers, in the transform's type
  if constexpr (IKX)
    for (int i = threadIdx.x; i <= M; i += row_block<T>()) ikxs[i] = (T)ikx[i];
  auto ld = [&](int k) -> cx<T> {
    cx<T> x = (valid && k < ncols_in) ? src[k] : mk<T>(0, 0);
    if (k == 0 || k == M) x.y = 0;
    return x;
  };
  __syncthreads();  // the twiddles
  cx<T> v[Q];
  if constexpr (IKX) {  // i kx X[k]: the kx = 0 and kx = M entries of the table are zero
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = t + P * q;
      const T sa = ikxs[k], sb = ikxs[M - k];
      const cx<T> ra = ld(k), rb = ld(M - k);
      v[q] = merge_c2r(mk<T>(-ra.y * sa, ra.x * sa), mk<T>(-rb.y * sb, rb.x * sb), tws[M + k]);
    }
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = t + P * q;
      v[q] = merge_c2r(ld(k), ld(M - k), tws[M + k]);
    }
  }
  F::stage1(v, t, sc, tws);
  slot_sync<P>();
  TO *dst = out + plane * out_plane + y * out_pitch;
  const TO *tzr = tz + plane * out_plane + y * out_pitch;
  const bool top = plane == nplanes - 1;
  const bool mid = plane >= 1 && plane <= nplanes - 2;
  const long up = top ? 0 : out_plane, dn = mid ? out_plane : 0;  // neighbour levels, clamped
  const double sh = (EPI == 0 && shift && valid) ? shift[plane] : 0.0;
  auto fin = [&](T v, int x) -> TO {
    if constexpr (EPI == 1) {
      const TO d = cst * (tzr[x + up] - tzr[x]);
      return top ? (TO)LP_BOGUS : (TO)v + d;
    } else if constexpr (EPI == 2) {
      TO r = (TO)v;
      const TO d = cst * (tzr[x] - tzr[x - dn]);
      if (mid) r += d;
      return r;
    } else {
      if (shift) return (TO)((double)v + sh);  // the plane's shift back, in binary64
      return (TO)v;
    }
  };
#pragma unroll
  for (int j = 0; j < F::G; ++j) {
    cx<T> a[P];
    if (F::stage2(j, t, sc, a) && valid) {
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        const int n = t + P * j + Q * k1;
        dst[2 * n] = fin(a[k1].x, 2 * n);
        dst[2 * n + 1] = fin(a[k1].y, 2 * n + 1);
      }
    }
  }
  if (valid) {
    const TO pad = (EPI == 1 && top) ? (TO)LP_BOGUS : (TO)0;
    for (long n = 2 * M + t; n < out_pitch; n += P) dst[n] = pad;
  }
}

// ---------------------------------------------------------------------------------
// col_op: per kx column, forward DFT along y, spectral map, inverse DFT along y.
// A block owns TX adjacent columns (one 128-byte segment per y) of one plane; 16 lanes
// per column.  The spectrum between the transforms stays in registers in the
// "D16" layout: lane t holds entries t + 16*m, which is also what the padding,
// truncation and multiplier maps need, so they are lane-local.
// ---------------------------------------------------------------------------------
struct ColSlot {
  void *ptr;
  long pitch;  // complex elements between y rows of this output
  long plane;  // complex elements between planes of this output
  int mult;
  int tpitch;  // row length of the multiplier tables (lh of the small grid)
  int ncols;   // columns stored (kx < ncols)
  double scale;  // factor on the padded / truncated spectrum (padmode 1, 2)
};
struct ColParams {
  ColSlot outs[3];
  int nout;
};
static constexpr int kColLanes = 16;
template <typename T> __host__ __device__ constexpr int col_tx() { return 128 / (int)sizeof(cx<T>); }
// LDS stride of a column slot.  The 16 lanes of a quarter wavefront are 16 (8 in binary64)
// different columns at the same row offset, so an odd stride puts them on different banks.
__host__ __device__ constexpr int col_slot_stride(int n) { return n | 1; }

// forward DFT of length NY along a column held in the D16 layout (sp[m] = x[t + 16 m]),
// result in the same layout; sc is the column's LDS slot (block-wide barriers)
template <typename T, int NY>
__device__ __forceinline__ void col_forward(cx<T> (&sp)[NY / kColLanes], int t, cx<T> *sc,
                                            const cx<T> *__restrict__ twI) {
  constexpr int PL = kColLanes;
  using FI = Fft4<T, NY, PL, -1>;
  constexpr int QI = FI::Q;
  FI::stage1(sp, t, sc, twI);
  __syncthreads();
  if constexpr (QI == PL) {  // stage 2 already yields the D16 layout
    cx<T> a[PL];
    FI::stage2(0, t, sc, a);
#pragma unroll
    for (int k1 = 0; k1 < PL; ++k1) sp[k1] = a[k1];
    __syncthreads();
  } else {  // regroup through LDS in natural order
    cx<T> a[FI::G][PL];
    bool ok[FI::G];
#pragma unroll
    for (int j = 0; j < FI::G; ++j) ok[j] = FI::stage2(j, t, sc, a[j]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < FI::G; ++j)
      if (ok[j]) {
#pragma unroll
        for (int k1 = 0; k1 < PL; ++k1) sc[t + PL * j + QI * k1] = a[j][k1];
      }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < QI; ++m) sp[m] = sc[t + PL * m];
    __syncthreads();
  }
}

// TX columns per block (16 for binary32, 8 for binary64: one 128-byte run per y); TX = 1
// serves a single trailing column.  Block column c handles kx = col0 + blockIdx.x*TX + c.
template <typename T, int NYI, int NYO, int TX = col_tx<T>()>
__global__ __launch_bounds__(TX * kColLanes) void k_col_op(
    const cx<T> *__restrict__ in, long in_pitch, long in_plane, int ncols, int fwd, int inv, ColParams Pr,
    long out_pitch, long out_plane, const double *__restrict__ kxt, const double *__restrict__ kyt,
    const double *__restrict__ G1, const double *__restrict__ G2, int lh_s, int ny_s,
    const cx<T> *__restrict__ twI, const cx<T> *__restrict__ twO, int col0, int nplanes) {
  constexpr int PL = kColLanes;
  using FI = Fft4<T, NYI, PL, -1>;
  using FO = Fft4<T, NYO, PL, +1>;
  constexpr int QI = FI::Q, QO = FO::Q;
  constexpr int PM = NYI == NYO ? 0 : (NYO > NYI ? 1 : 2);  // same grid / pad / truncate
  constexpr int DQ = (NYI > NYO ? NYI - NYO : NYO - NYI) / PL;
  constexpr int NS = FI::SCR > FO::SCR ? (FI::SCR > NYI ? FI::SCR : NYI) : (FO::SCR > NYI ? FO::SCR : NYI);
  constexpr int SS = col_slot_stride(NS);
  // binary32: the twiddles are read from LDS (staged once per workgroup), which also
  // lowers the register count; binary64 keeps the slots' LDS to itself and reads the tables
  constexpr bool TWL = sizeof(T) == 4;
  __shared__ cx<T> lds[TX * SS + (TWL ? NYI + NYO : 0)];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  int plane = blockIdx.y, cg = blockIdx.x;
  if constexpr (xcd_order<T>()) {  // last-written planes first; the grid has gx x 8 max_j(planes of j) blocks
    const int gx = gridDim.x;
    const long lin = blockIdx.x + (long)gx * blockIdx.y;
    const int xj = (int)(lin & 7), kq = (int)((lin >> 3) / gx);
    const int npj = xcd_planes(nplanes, xj);
    if (kq >= npj) return;  // uniform over the block
    plane = xj + 8 * (npj - 1 - kq);
    cg = (int)((lin >> 3) % gx);
  }
  const int kx = col0 + cg * TX + c;
  const bool cv = kx < ncols;
  cx<T> *sc = lds + c * SS;
  if constexpr (TWL) {
    cx<T> *tl = lds + TX * SS;
    for (int i = threadIdx.x; i < NYI + NYO; i += TX * kColLanes) tl[i] = i < NYI ? twI[i] : twO[i - NYI];
    twI = tl;
    twO = tl + NYI;
    __syncthreads();
  }
  // invalid columns load column ncols-1 and discard it (no divergent loads)
  const cx<T> *src = in + (long)plane * in_plane + (cv ? kx : ncols - 1) + (long)t * in_pitch;
  const long istep = (long)PL * in_pitch;
  cx<T> sp[QI];  // D16: sp[m] = spectrum[t + 16 m]
#pragma unroll
  for (int m = 0; m < QI; ++m) {
    const cx<T> x = src[m * istep];
    sp[m] = cv ? x : mk<T>(0, 0);
  }
  if (fwd) col_forward<T, NYI>(sp, t, sc, twI);
  for (int jo = 0; jo < Pr.nout; ++jo) {
    const int mult = Pr.outs[jo].mult;
    // one multiplier table per output: i*kx, i*ky (imaginary) or a filter (real)
    const double *mt = mult == M_IKX ? kxt : mult == M_IKY ? kyt : mult == M_G1 ? G1 : G2;
    const bool imag = mult == M_IKX || mult == M_IKY;
    const bool keep = cv && kx != lh_s - 1;
    const int tp = Pr.outs[jo].tpitch;
    const double *mrow = mt + (long)t * tp + (cv ? kx : 0);
    cx<T> v[QO];
    if constexpr (PM == 0) {
      // the multipliers of all q are loaded together (block-uniform branch)
      T fq[QO];
      if (mult != M_ONE) {
#pragma unroll
        for (int q = 0; q < QO; ++q) fq[q] = (T)mrow[(long)q * PL * tp];
      }
#pragma unroll
      for (int q = 0; q < QO; ++q) {
        const int yo = t + PL * q;
        cx<T> x = sp[q];
        if (mult != M_ONE) x = mk<T>((imag ? -x.y : x.x) * fq[q], (imag ? x.x : x.y) * fq[q]);
        v[q] = (!keep || yo == ny_s / 2) ? mk<T>(0, 0) : x;
      }
    }
    if constexpr (PM != 0) {
      const T osc = (T)Pr.outs[jo].scale;
      const bool scaled = Pr.outs[jo].scale != 1.0;
#pragma unroll
      for (int q = 0; q < QO; ++q) {
        const int yo = t + PL * q;
        cx<T> x = mk<T>(0, 0);
        if constexpr (PM == 1) {  // padd.f90: ny -> ny2, middle rows zero
          if (q < QI && yo < NYI / 2) x = sp[q < QI ? q : 0];
          else if (q >= DQ && yo >= NYO - NYI / 2 + 1) x = sp[q >= DQ ? q - DQ : 0];
        } else {  // unpadd.f90: ny2 -> ny, Nyquist row zero
          if (yo < NYO / 2) x = sp[q];
          else if (yo > NYO / 2) x = sp[q + DQ];
        }
        v[q] = scaled ? mk<T>(x.x * osc, x.y * osc) : x;
      }
    }
    // addresses are formed from per-output values so that they are not hoisted out of
    // the output loop (which would keep 16 64-bit addresses live)
    const long pitch = Pr.outs[jo].pitch;
    const bool st = cv && kx < Pr.outs[jo].ncols;
    cx<T> *dst = reinterpret_cast<cx<T> *>(Pr.outs[jo].ptr) + (long)plane * Pr.outs[jo].plane + kx;
    if (inv) {
      FO::stage1(v, t, sc, twO);
      __syncthreads();
#pragma unroll
      for (int j = 0; j < FO::G; ++j) {
        cx<T> a[PL];
        if (FO::stage2(j, t, sc, a) && st) {
          cx<T> *d = dst + (long)(t + PL * j) * pitch;
          const long ostep = (long)QO * pitch;
#pragma unroll
          for (int k1 = 0; k1 < PL; ++k1) d[k1 * ostep] = a[k1];
        }
      }
      __syncthreads();
    } else if (st) {
      cx<T> *d = dst + (long)t * pitch;
      const long ostep = (long)PL * pitch;
#pragma unroll
      for (int q = 0; q < QO; ++q) d[q * ostep] = v[q];
    }
  }
}

// ---------------------------------------------------------------------------------
// col_div: the x- and y-derivative terms of divstress_uv.f90 / divstress_w.f90 summed in
// spectral space.  Per kx column: forward y DFTs of the x-spectra A and B, then
// i kx A + i ky B with the Nyquist row and column zeroed (the ddx / ddy multipliers),
// inverse y DFT into out.
// ---------------------------------------------------------------------------------
template <typename T, int NY>
__global__ __launch_bounds__(col_tx<T>() * kColLanes) void k_col_div(
    const cx<T> *__restrict__ A, const cx<T> *__restrict__ B, long pitch, long plane_el, int ncols,
    cx<T> *__restrict__ out, const double *__restrict__ kxt, const double *__restrict__ kyt, int lh_s, int ny_s,
    const cx<T> *__restrict__ tw) {
  constexpr int TX = col_tx<T>();
  constexpr int PL = kColLanes;
  using FO = Fft4<T, NY, PL, +1>;
  constexpr int Q = NY / PL;
  constexpr int SS = col_slot_stride(FO::SCR > NY ? FO::SCR : NY);
  constexpr bool TWL = (TX * SS + NY) * sizeof(cx<T>) <= 40960;  // as in k_col_op
  __shared__ cx<T> lds[TX * SS + (TWL ? NY : 0)];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int kx = blockIdx.x * TX + c;
  const bool cv = kx < ncols;
  const long plane = blockIdx.y;
  cx<T> *sc = lds + c * SS;
  if constexpr (TWL) {
    cx<T> *tl = lds + TX * SS;
    for (int i = threadIdx.x; i < NY; i += TX * kColLanes) tl[i] = tw[i];
    tw = tl;
    __syncthreads();
  }
  const long off = plane * plane_el + (cv ? kx : ncols - 1) + (long)t * pitch;
  const long istep = (long)PL * pitch;
  cx<T> sa[Q], sb[Q];  // D16
#pragma unroll
  for (int m = 0; m < Q; ++m) {
    sa[m] = A[off + m * istep];
    sb[m] = B[off + m * istep];
  }
  col_forward<T, NY>(sa, t, sc, tw);
  col_forward<T, NY>(sb, t, sc, tw);
  const bool keep = cv && kx != lh_s - 1;
  const long mo = (long)t * lh_s + (cv ? kx : 0);
  T fx[Q], fy[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    fx[q] = (T)kxt[mo + (long)q * PL * lh_s];
    fy[q] = (T)kyt[mo + (long)q * PL * lh_s];
  }
  cx<T> v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int yo = t + PL * q;
    const cx<T> a = sa[q], b = sb[q];
    const cx<T> x = mk<T>(-a.y * fx[q] + -b.y * fy[q], a.x * fx[q] + b.x * fy[q]);
    v[q] = (!keep || yo == ny_s / 2) ? mk<T>(0, 0) : x;
  }
  FO::stage1(v, t, sc, tw);
  __syncthreads();
  cx<T> *dst = out + plane * plane_el + kx;
#pragma unroll
  for (int j = 0; j < FO::G; ++j) {
    cx<T> a[PL];
    if (FO::stage2(j, t, sc, a) && cv) {
      cx<T> *d = dst + (long)(t + PL * j) * pitch;
      const long ostep = (long)Q * pitch;
#pragma unroll
      for (int k1 = 0; k1 < PL; ++k1) d[k1 * ostep] = a[k1];
    }
  }
}

template <typename T, int NY>
static hipError_t launch_col_div(const void *A, const void *B, long pitch, long plane, int ncols, int nplanes,
                                 void *out, const SpecTables &tab, hipStream_t s, Dt dt) {
  const void *tw = tw_ptr(NY, dt);
  if (!tw || ncols < 1 || ncols > tab.lh || pitch < ncols || tab.ny != NY) return hipErrorInvalidValue;
  constexpr int TX = col_tx<T>();
  dim3 grid((unsigned)((ncols + TX - 1) / TX), (unsigned)nplanes);
  hipLaunchKernelGGL((k_col_div<T, NY>), grid, dim3(TX * kColLanes), 0, s, (const cx<T> *)A, (const cx<T> *)B,
                     pitch, plane, ncols, (cx<T> *)out, tab.kx, tab.ky, tab.lh, tab.ny, (const cx<T> *)tw);
  return hipGetLastError();
}

hipError_t col_div(Dt dt, int ny, const void *A, const void *B, long pitch, long plane, int ncols, int nplanes,
                   void *out, const SpecTables &tab, hipStream_t s) {
#define CASE(n)                                                                                      \
  case n:                                                                                            \
    return dt == F64 ? launch_col_div<double, n>(A, B, pitch, plane, ncols, nplanes, out, tab, s, dt) \
                     : launch_col_div<float, n>(A, B, pitch, plane, ncols, nplanes, out, tab, s, dt);
  switch (ny) {
    CASE(32) CASE(48) CASE(64) CASE(96) CASE(128) CASE(144) CASE(192) CASE(256)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

// ---------------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------------
#define LP_ROW_LENGTHS(X) X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192)
#define LP_SMALL_ROW_LENGTHS(X) X(16) X(32) X(48) X(64) X(96) X(128)

bool fft_len_supported(int n) {
  switch (n) {
    case 32: case 48: case 64: case 96: case 128: case 144: case 192: case 256:
    case 288: case 384: return true;
    default: return false;
  }
}

template <typename T, int M, int PRO = 0, typename TI = T, bool SH = false>
static hipError_t launch_r2c(const void *in, long in_pitch, long in_plane, int rows, int nplanes,
                             double scale, void *out, long out_pitch, long out_plane, int ncols_out,
                             hipStream_t s, Dt dt, const void *pa = nullptr, const void *pb = nullptr,
                             double *shift_out = nullptr) {
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN) return hipErrorInvalidValue;
  constexpr int RPB = kR2cBlock / row_lanes(M);
  long total = (long)rows * nplanes;
  if (total < 1) return hipErrorInvalidValue;
  dim3 grid(r2c_grid((total + RPB - 1) / RPB, r2c_persistent<T>()));
  hipLaunchKernelGGL((k_row_r2c<T, M, PRO, TI, SH>), grid, dim3(kR2cBlock), 0, s, (const TI *)in, in_pitch,
                     in_plane, rows, nplanes, (TI)scale, (cx<T> *)out, out_pitch, out_plane, ncols_out,
                     (const cx<T> *)twM, (const cx<T> *)twN, (const T *)pa, (const T *)pb, shift_out);
  return hipGetLastError();
}

hipError_t row_r2c_vort(Dt dt, int M, int comp, const void *a, const void *b, long in_pitch, long in_plane,
                        int rows, int nplanes, double scale, void *out, long out_pitch, long out_plane,
                        int ncols_out, hipStream_t s) {
  if (!a || !b || comp < 0 || comp > 2 || nplanes < 2) return hipErrorInvalidValue;
#define CASE(m)                                                                                           \
  case m:                                                                                                 \
    if (dt == F64) {                                                                                      \
      if (comp == 0) return launch_r2c<double, m, 1>(nullptr, in_pitch, in_plane, rows, nplanes, scale,   \
                                                     out, out_pitch, out_plane, ncols_out, s, dt, a, b);  \
      if (comp == 1) return launch_r2c<double, m, 2>(nullptr, in_pitch, in_plane, rows, nplanes, scale,   \
                                                     out, out_pitch, out_plane, ncols_out, s, dt, a, b);  \
      return launch_r2c<double, m, 3>(nullptr, in_pitch, in_plane, rows, nplanes, scale, out, out_pitch,  \
                                      out_plane, ncols_out, s, dt, a, b);                                 \
    }                                                                                                     \
    if (comp == 0) return launch_r2c<float, m, 1>(nullptr, in_pitch, in_plane, rows, nplanes, scale, out, \
                                                  out_pitch, out_plane, ncols_out, s, dt, a, b);          \
    if (comp == 1) return launch_r2c<float, m, 2>(nullptr, in_pitch, in_plane, rows, nplanes, scale, out, \
                                                  out_pitch, out_plane, ncols_out, s, dt, a, b);          \
    return launch_r2c<float, m, 3>(nullptr, in_pitch, in_plane, rows, nplanes, scale, out, out_pitch,     \
                                   out_plane, ncols_out, s, dt, a, b);
  switch (M) {
    LP_SMALL_ROW_LENGTHS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

hipError_t row_r2c(Dt dt, int M, const void *in, long in_pitch, long in_plane, int rows, int nplanes,
                   double scale, void *out, long out_pitch, long out_plane, int ncols_out, hipStream_t s,
                   RowIn ri) {
  const Dt din = ri.din < 0 ? dt : (Dt)ri.din;
  if (ri.shift && !(dt == F32 && din == F64)) return hipErrorInvalidValue;
  if (dt == F64 && din == F32) return hipErrorInvalidValue;
#define CASE(m)                                                                                        \
  case m:                                                                                              \
    if (dt == F64)                                                                                     \
      return launch_r2c<double, m>(in, in_pitch, in_plane, rows, nplanes, scale, out, out_pitch,       \
                                   out_plane, ncols_out, s, dt);                                       \
    if (din == F32)                                                                                    \
      return launch_r2c<float, m>(in, in_pitch, in_plane, rows, nplanes, scale, out, out_pitch,        \
                                  out_plane, ncols_out, s, dt);                                        \
    if (ri.shift)                                                                                      \
      return launch_r2c<float, m, 0, double, true>(in, in_pitch, in_plane, rows, nplanes, scale, out,  \
                                                   out_pitch, out_plane, ncols_out, s, dt, nullptr,    \
                                                   nullptr, ri.shift_out);                             \
    return launch_r2c<float, m, 0, double, false>(in, in_pitch, in_plane, rows, nplanes, scale, out,   \
                                                  out_pitch, out_plane, ncols_out, s, dt);
  switch (M) {
    LP_ROW_LENGTHS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

template <typename T, typename TO, int M, int EPI = 0>
static hipError_t launch_c2r(const void *in, long in_pitch, long in_plane, int ncols_in, int rows,
                             int nplanes, void *out, long out_pitch, long out_plane, hipStream_t s, Dt dt,
                             const void *tz = nullptr, double cst = 0.0, const double *ikx = nullptr,
                             const double *shift = nullptr) {
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN) return hipErrorInvalidValue;
  constexpr int RPB = row_block<T>() / row_lanes(M);
  long total = (long)rows * nplanes;
  const int xorder = xcd_order<T>() && rows % RPB == 0 ? 1 : 0;
  dim3 grid((unsigned)(xorder ? 8L * xcd_planes(nplanes, 0) * (rows / RPB) : (total + RPB - 1) / RPB));
  // the x derivative is its own kernel: the plain transform then holds no multiplier
  // registers
  if (ikx) {
    if constexpr (EPI != 0) return hipErrorInvalidValue;  // no caller asks for both
    else
      hipLaunchKernelGGL((k_row_c2r<T, TO, M, EPI, true>), grid, dim3(row_block<T>()), 0, s, (const cx<T> *)in,
                         in_pitch, in_plane, ncols_in, rows, nplanes, (TO *)out, out_pitch, out_plane,
                         (const cx<T> *)twM, (const cx<T> *)twN, (const TO *)out, (TO)cst, ikx, shift, xorder);
  } else
    hipLaunchKernelGGL((k_row_c2r<T, TO, M, EPI, false>), grid, dim3(row_block<T>()), 0, s, (const cx<T> *)in,
                       in_pitch, in_plane, ncols_in, rows, nplanes, (TO *)out, out_pitch, out_plane,
                       (const cx<T> *)twM, (const cx<T> *)twN, (const TO *)(EPI ? tz : out), (TO)cst, nullptr,
                       shift, xorder);
  return hipGetLastError();
}

hipError_t row_c2r_div(Dt dt, int M, const void *in, long in_pitch, long in_plane, int ncols_in, int rows,
                       int nplanes, void *out, long out_pitch, long out_plane, const void *tz, double cst,
                       int is_w, hipStream_t s) {
  if (!tz || nplanes < 2 || out_pitch < 2 * M) return hipErrorInvalidValue;
#define CASE(m)                                                                                          \
  case m:                                                                                                \
    if (dt == F64)                                                                                       \
      return is_w ? launch_c2r<double, double, m, 2>(in, in_pitch, in_plane, ncols_in, rows, nplanes,    \
                                                     out, out_pitch, out_plane, s, dt, tz, cst)          \
                  : launch_c2r<double, double, m, 1>(in, in_pitch, in_plane, ncols_in, rows, nplanes,    \
                                                     out, out_pitch, out_plane, s, dt, tz, cst);         \
    return is_w ? launch_c2r<float, float, m, 2>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,   \
                                                 out_pitch, out_plane, s, dt, tz, cst)                   \
                : launch_c2r<float, float, m, 1>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,   \
                                                 out_pitch, out_plane, s, dt, tz, cst);
  switch (M) {
    LP_SMALL_ROW_LENGTHS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

hipError_t row_c2r(Dt dtc, Dt dto, int M, const void *in, long in_pitch, long in_plane, int ncols_in,
                   int rows, int nplanes, void *out, long out_pitch, long out_plane, hipStream_t s,
                   const double *ikx, const double *shift) {
  if (shift && !(dtc == F32 && dto == F64)) return hipErrorInvalidValue;
#define CASE(m)                                                                                         \
  case m:                                                                                               \
    if (dtc == F64 && dto == F64)                                                                       \
      return launch_c2r<double, double, m>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,        \
                                           out_pitch, out_plane, s, dtc, nullptr, 0.0, ikx);            \
    if (dtc == F64 && dto == F32)                                                                       \
      return launch_c2r<double, float, m>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,         \
                                          out_pitch, out_plane, s, dtc, nullptr, 0.0, ikx);             \
    if (dtc == F32 && dto == F32)                                                                       \
      return launch_c2r<float, float, m>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,          \
                                         out_pitch, out_plane, s, dtc, nullptr, 0.0, ikx);              \
    if (dtc == F32 && dto == F64)                                                                       \
      return launch_c2r<float, double, m>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,         \
                                          out_pitch, out_plane, s, dtc, nullptr, 0.0, ikx, shift);      \
    return hipErrorInvalidValue;
  switch (M) {
    LP_ROW_LENGTHS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

// col0 > 0: the columns col0..ncols-1 only (same grid, one column per block)
template <typename T, int NYI, int NYO>
static hipError_t launch_col(int padmode, int fwd, int inv, const void *in, long in_pitch, long in_plane,
                             int ncols, int nplanes, int nout, const ColOut *outs, long out_pitch,
                             long out_plane, const SpecTables &tab, hipStream_t s, Dt dt, int col0 = 0,
                             LaunchEv ev = {}) {
  const void *twI = tw_ptr(NYI, dt), *twO = tw_ptr(NYO, dt);
  if (!twI || !twO || col0 < 0 || col0 >= ncols) return hipErrorInvalidValue;
  // the map is fixed by the lengths: same grid (0), padd (1), unpadd (2)
  if (padmode != (NYI == NYO ? 0 : (NYO > NYI ? 1 : 2))) return hipErrorInvalidValue;
  constexpr int TX = col_tx<T>();
  ColParams P;
  P.nout = nout;
  for (int j = 0; j < 3; ++j) {
    if (j >= nout) {
      P.outs[j] = ColSlot{nullptr, 0, 0, 0, 0, 0, 1.0};
      continue;
    }
    const ColOut &o = outs[j];
    if (padmode == 0 && o.scale != 1.0) return hipErrorInvalidValue;
    P.outs[j] = ColSlot{o.ptr, o.pitch ? o.pitch : out_pitch, o.plane ? o.plane : out_plane, o.mult, tab.lh,
                        o.ncols ? o.ncols : ncols, o.scale};
  }
  if (col0 > 0) {
    if constexpr (NYI == NYO) {
      dim3 grid((unsigned)(ncols - col0), (unsigned)(xcd_order<T>() ? 8 * xcd_planes(nplanes, 0) : nplanes));
      hipExtLaunchKernelGGL((k_col_op<T, NYI, NYO, 1>), grid, dim3(kColLanes), 0, s, ev.start, ev.stop, 0,
                            (const cx<T> *)in, in_pitch, in_plane, ncols, fwd, inv, P, out_pitch, out_plane,
                            tab.kx, tab.ky, tab.G1, tab.G2, tab.lh, tab.ny, (const cx<T> *)twI,
                            (const cx<T> *)twO, col0, nplanes);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  dim3 grid((unsigned)((ncols + TX - 1) / TX),
            (unsigned)(xcd_order<T>() ? 8 * xcd_planes(nplanes, 0) : nplanes));
  hipExtLaunchKernelGGL((k_col_op<T, NYI, NYO>), grid, dim3(TX * kColLanes), 0, s, ev.start, ev.stop, 0,
                        (const cx<T> *)in, in_pitch, in_plane, ncols, fwd, inv, P, out_pitch, out_plane, tab.kx,
                        tab.ky, tab.G1, tab.G2, tab.lh, tab.ny, (const cx<T> *)twI, (const cx<T> *)twO, 0,
                        nplanes);
  return hipGetLastError();
}

hipError_t col_op_tail(Dt dt, int ny, const void *in, long in_pitch, long in_plane, int col0, int ncols,
                       int nplanes, int nout, const ColOut *outs, long out_pitch, long out_plane,
                       const SpecTables &tab, hipStream_t s, LaunchEv ev) {
  if (nout < 1 || nout > 3) return hipErrorInvalidValue;
#define CASE(n)                                                                                          \
  case n:                                                                                                \
    return dt == F64 ? launch_col<double, n, n>(0, 1, 1, in, in_pitch, in_plane, ncols, nplanes, nout,  \
                                                outs, out_pitch, out_plane, tab, s, dt, col0, ev)       \
                     : launch_col<float, n, n>(0, 1, 1, in, in_pitch, in_plane, ncols, nplanes, nout,   \
                                               outs, out_pitch, out_plane, tab, s, dt, col0, ev);
  switch (ny) {
    CASE(32) CASE(48) CASE(64) CASE(96) CASE(128) CASE(144) CASE(192) CASE(256) CASE(288) CASE(384)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

hipError_t col_op(Dt dt, int nyi, int nyo, int padmode, int fwd, int inv, const void *in, long in_pitch,
                  long in_plane, int ncols, int nplanes, int nout, const ColOut *outs, long out_pitch,
                  long out_plane, const SpecTables &tab, hipStream_t s, LaunchEv ev) {
  if (nout < 1 || nout > 3) return hipErrorInvalidValue;
#define SAME(n)                                                                                         \
  if (nyi == n && nyo == n)                                                                             \
    return dt == F64 ? launch_col<double, n, n>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,       \
                                                nplanes, nout, outs, out_pitch, out_plane, tab, s, dt, 0, ev) \
                     : launch_col<float, n, n>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,        \
                                               nplanes, nout, outs, out_pitch, out_plane, tab, s, dt, 0, ev);
#define PAIR(a, b)                                                                                      \
  if (nyi == a && nyo == b)                                                                             \
    return dt == F64 ? launch_col<double, a, b>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,       \
                                                nplanes, nout, outs, out_pitch, out_plane, tab, s, dt, 0, ev) \
                     : launch_col<float, a, b>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,        \
                                               nplanes, nout, outs, out_pitch, out_plane, tab, s, dt, 0, ev);
  SAME(32) SAME(48) SAME(64) SAME(96) SAME(128) SAME(144) SAME(192) SAME(256) SAME(288) SAME(384)
  PAIR(32, 48) PAIR(48, 32) PAIR(64, 96) PAIR(96, 64) PAIR(96, 144) PAIR(144, 96)
  PAIR(128, 192) PAIR(192, 128) PAIR(192, 288) PAIR(288, 192) PAIR(256, 384) PAIR(384, 256)
#undef SAME
#undef PAIR
  return hipErrorInvalidValue;
}

}  // namespace lp
