// lp_fft.hip -- batched plane FFT passes for gfx950.
//
// A 2-D real transform of an (ld,ny) plane (FFTW2 rfftwnd in-place layout,
// fft.f90:35-46) is split into
//   row_r2c : each x-row of 2M reals -> M+1 complex (half-length complex FFT + split)
//   col_op  : per kx column, complex FFT along y, a spectral map, inverse FFT along y
//   row_c2r : M+1 complex -> 2M reals (merge + half-length inverse FFT), using only the
//             real parts of the kx = 0 and kx = M terms (FFTW2 halfcomplex semantics)
// Rows and columns are staged in LDS; each wave64 group handles whole transforms.
#include "lp_device.hpp"
#include "lp_internal.hpp"

namespace lp {

static constexpr int kBlock = 256;

template <typename T> struct V2;
template <> struct V2<double> { using t = double2; };
template <> struct V2<float> { using t = float2; };

__host__ __device__ constexpr int row_tpr(int M) {  // threads per row
  return M >= 128 ? 64 : M >= 64 ? 32 : M >= 32 ? 16 : 8;
}

// ---------------------------------------------------------------------------------
// row_r2c
// ---------------------------------------------------------------------------------
template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_row_r2c(const T *__restrict__ in, long in_pitch, long in_plane,
                                                    int rows, int nplanes, T scale, cx<T> *__restrict__ out,
                                                    long out_pitch, long out_plane, int ncols_out,
                                                    const cx<T> *__restrict__ twM,
                                                    const cx<T> *__restrict__ twN) {
  constexpr int TPR = row_tpr(M);
  constexpr int RPB = kBlock / TPR;
  __shared__ cx<T> sA[RPB][M];
  __shared__ cx<T> sB[RPB][M];
  const int r = threadIdx.x / TPR, t = threadIdx.x % TPR;
  const long grow = (long)blockIdx.x * RPB + r;
  const long total = (long)rows * nplanes;
  const bool valid = grow < total;
  const long plane = valid ? grow / rows : 0, y = valid ? grow % rows : 0;
  const T *src = in + plane * in_plane + y * in_pitch;
  using v2 = typename V2<T>::t;
  if (valid) {
    for (int n = t; n < M; n += TPR) {
      v2 p = *reinterpret_cast<const v2 *>(src + 2 * n);
      sA[r][n] = mk<T>(p.x * scale, p.y * scale);
    }
  }
  __syncthreads();
  cx<T> *Z = fft_lds<T, M, -1>(&sA[r][0], &sB[r][0], t, TPR, twM);
  // split: X[k] = (Z[k]+conj(Z[M-k]))/2 - i/2 w^k (Z[k]-conj(Z[M-k])), w = exp(-2 pi i/2M)
  if (valid) {
    cx<T> *dst = out + plane * out_plane + y * out_pitch;
    for (int k = t; k <= M && k < ncols_out; k += TPR) {
      cx<T> a = Z[k % M], b = conjg(Z[(M - k) % M]);
      cx<T> e = a + b, o = a - b;
      cx<T> w = twN[k % (2 * M)];
      cx<T> wo = cmul(w, o);
      // -i * wo
      cx<T> X = mk<T>(T(0.5) * (e.x + wo.y), T(0.5) * (e.y - wo.x));
      dst[k] = X;
    }
  }
}

// ---------------------------------------------------------------------------------
// row_c2r
// ---------------------------------------------------------------------------------
template <typename T, typename TO, int M>
__global__ __launch_bounds__(kBlock) void k_row_c2r(const cx<T> *__restrict__ in, long in_pitch, long in_plane,
                                                    int ncols_in, int rows, int nplanes,
                                                    TO *__restrict__ out, long out_pitch, long out_plane,
                                                    const cx<T> *__restrict__ twM,
                                                    const cx<T> *__restrict__ twN) {
  constexpr int TPR = row_tpr(M);
  constexpr int RPB = kBlock / TPR;
  __shared__ cx<T> sX[RPB][M + 1];
  __shared__ cx<T> sA[RPB][M];
  __shared__ cx<T> sB[RPB][M];
  const int r = threadIdx.x / TPR, t = threadIdx.x % TPR;
  const long grow = (long)blockIdx.x * RPB + r;
  const long total = (long)rows * nplanes;
  const bool valid = grow < total;
  const long plane = valid ? grow / rows : 0, y = valid ? grow % rows : 0;
  if (valid) {
    const cx<T> *src = in + plane * in_plane + y * in_pitch;
    for (int k = t; k <= M; k += TPR) {
      cx<T> v = k < ncols_in ? src[k] : mk<T>(0, 0);
      if (k == 0 || k == M) v.y = 0;
      sX[r][k] = v;
    }
  }
  __syncthreads();
  // merge: Y[k] = (X[k]+conj(X[M-k])) + i w^-k (X[k]-conj(X[M-k]))
  for (int k = t; k < M; k += TPR) {
    cx<T> a = sX[r][k], b = conjg(sX[r][M - k]);
    cx<T> e = a + b, o = a - b;
    cx<T> w = conjg(twN[k]);
    cx<T> wo = cmul(w, o);
    sA[r][k] = mk<T>(e.x - wo.y, e.y + wo.x);
  }
  __syncthreads();
  cx<T> *Yv = fft_lds<T, M, +1>(&sA[r][0], &sB[r][0], t, TPR, twM);
  if (valid) {
    TO *dst = out + plane * out_plane + y * out_pitch;
    for (int n = t; n < M; n += TPR) {
      cx<T> v = Yv[n];
      dst[2 * n] = (TO)v.x;
      dst[2 * n + 1] = (TO)v.y;
    }
    for (long n = 2 * M + t; n < out_pitch; n += TPR) dst[n] = (TO)0;
  }
}

// ---------------------------------------------------------------------------------
// col_op
// ---------------------------------------------------------------------------------
struct ColParams {
  ColOut outs[3];
  int nout;
};

__host__ __device__ constexpr int col_tx(int nmax, int bytes_cx) {
  // columns per block: 3 buffers of nmax complex per column within 96 KiB
  return (3 * nmax * bytes_cx * 16 <= 98304) ? 16
         : (3 * nmax * bytes_cx * 8 <= 98304) ? 8
         : (3 * nmax * bytes_cx * 4 <= 98304) ? 4
         : (3 * nmax * bytes_cx * 2 <= 98304) ? 2 : 1;
}

template <typename T, int NYI, int NYO>
__global__ __launch_bounds__(kBlock) void k_col_op(const cx<T> *__restrict__ in, long in_pitch, long in_plane,
                                                   int ncols, int nplanes, int padmode, int fwd, int inv,
                                                   ColParams P, long out_pitch, long out_plane,
                                                   const double *__restrict__ kxt, const double *__restrict__ kyt,
                                                   const double *__restrict__ G1, const double *__restrict__ G2,
                                                   int lh_s, int ny_s,
                                                   const cx<T> *__restrict__ twI, const cx<T> *__restrict__ twO) {
  constexpr int NMAX = NYI > NYO ? NYI : NYO;
  constexpr int TX = col_tx(NMAX, (int)sizeof(cx<T>));
  constexpr int TPC = kBlock / TX;
  __shared__ cx<T> sbuf[3][TX][NMAX];
  const int c = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int plane = blockIdx.y;
  const int kx0 = blockIdx.x * TX;
  const cx<T> *src = in + (long)plane * in_plane;
  // coalesced load: element i -> (y = i / TX, col = i % TX)
  for (int i = threadIdx.x; i < NYI * TX; i += kBlock) {
    int yy = i / TX, cc = i % TX;
    int kx = kx0 + cc;
    cx<T> v = mk<T>(0, 0);
    if (kx < ncols) v = src[(long)yy * in_pitch + kx];
    sbuf[0][cc][yy] = v;
  }
  __syncthreads();
  int ib = 0;  // buffer holding the input spectrum
  if (fwd) {
    cx<T> *res = fft_lds<T, NYI, -1>(&sbuf[0][c][0], &sbuf[1][c][0], t, TPC, twI);
    ib = (res == &sbuf[0][c][0]) ? 0 : 1;
  }
  // ib is identical for every column (same stage count), so it is block-uniform
  const int b1 = (ib == 0) ? 1 : 0, b2 = 2;
  const int kx = kx0 + c;
  for (int j = 0; j < P.nout; ++j) {
    __syncthreads();
    const int mult = P.outs[j].mult;
    for (int yo = t; yo < NYO; yo += TPC) {
      cx<T> v = mk<T>(0, 0);
      int yi;
      if (padmode == 0) {
        yi = yo;
        if (kx < ncols && yo != ny_s / 2 && kx != lh_s - 1) {
          v = sbuf[ib][c][yi];
          long ti = (long)yo * lh_s + kx;
          if (mult == M_IKX) v = mul_i(v, (T)kxt[ti]);
          else if (mult == M_IKY) v = mul_i(v, (T)kyt[ti]);
          else if (mult == M_G1) v = scal(v, (T)G1[ti]);
          else if (mult == M_G2) v = scal(v, (T)G2[ti]);
        }
      } else if (padmode == 1) {  // small (NYI = ny) -> big (NYO = ny2)
        if (yo < NYI / 2) yi = yo;
        else if (yo >= NYO - NYI / 2 + 1) yi = yo - NYO + NYI;
        else yi = -1;
        if (yi >= 0 && kx < ncols) v = sbuf[ib][c][yi];
      } else {  // big (NYI = ny2) -> small (NYO = ny)
        if (yo < NYO / 2) yi = yo;
        else if (yo > NYO / 2) yi = yo + NYI - NYO;
        else yi = -1;
        if (yi >= 0 && kx < ncols) v = sbuf[ib][c][yi];
      }
      sbuf[b1][c][yo] = v;
    }
    __syncthreads();
    cx<T> *res = &sbuf[b1][c][0];
    if (inv) res = fft_lds<T, NYO, +1>(&sbuf[b1][c][0], &sbuf[b2][c][0], t, TPC, twO);
    __syncthreads();
    // coalesced store
    cx<T> *dst = reinterpret_cast<cx<T> *>(P.outs[j].ptr) + (long)plane * out_plane;
    const int rb = (res == &sbuf[b1][c][0]) ? b1 : b2;
    for (int i = threadIdx.x; i < NYO * TX; i += kBlock) {
      int yy = i / TX, cc = i % TX;
      int kxx = kx0 + cc;
      if (kxx < ncols) dst[(long)yy * out_pitch + kxx] = sbuf[rb][cc][yy];
    }
  }
}

// ---------------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------------
#define LP_ROW_LENGTHS(X) X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192)

bool fft_len_supported(int n) {
  switch (n) {
    case 32: case 48: case 64: case 96: case 128: case 144: case 192: case 256:
    case 288: case 384: return true;
    default: return false;
  }
}

template <typename T, int M>
static hipError_t launch_r2c(const void *in, long in_pitch, long in_plane, int rows, int nplanes,
                             double scale, void *out, long out_pitch, long out_plane, int ncols_out,
                             hipStream_t s, Dt dt) {
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN) return hipErrorInvalidValue;
  constexpr int RPB = kBlock / row_tpr(M);
  long total = (long)rows * nplanes;
  dim3 grid((unsigned)((total + RPB - 1) / RPB));
  hipLaunchKernelGGL((k_row_r2c<T, M>), grid, dim3(kBlock), 0, s, (const T *)in, in_pitch, in_plane, rows,
                     nplanes, (T)scale, (cx<T> *)out, out_pitch, out_plane, ncols_out,
                     (const cx<T> *)twM, (const cx<T> *)twN);
  return hipGetLastError();
}

hipError_t row_r2c(Dt dt, int M, const void *in, long in_pitch, long in_plane, int rows, int nplanes,
                   double scale, void *out, long out_pitch, long out_plane, int ncols_out, hipStream_t s) {
#define CASE(m)                                                                                        \
  case m:                                                                                              \
    return dt == F64 ? launch_r2c<double, m>(in, in_pitch, in_plane, rows, nplanes, scale, out,        \
                                             out_pitch, out_plane, ncols_out, s, dt)                   \
                     : launch_r2c<float, m>(in, in_pitch, in_plane, rows, nplanes, scale, out,         \
                                            out_pitch, out_plane, ncols_out, s, dt);
  switch (M) {
    LP_ROW_LENGTHS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

template <typename T, typename TO, int M>
static hipError_t launch_c2r(const void *in, long in_pitch, long in_plane, int ncols_in, int rows,
                             int nplanes, void *out, long out_pitch, long out_plane, hipStream_t s, Dt dt) {
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN) return hipErrorInvalidValue;
  constexpr int RPB = kBlock / row_tpr(M);
  long total = (long)rows * nplanes;
  dim3 grid((unsigned)((total + RPB - 1) / RPB));
  hipLaunchKernelGGL((k_row_c2r<T, TO, M>), grid, dim3(kBlock), 0, s, (const cx<T> *)in, in_pitch, in_plane,
                     ncols_in, rows, nplanes, (TO *)out, out_pitch, out_plane, (const cx<T> *)twM,
                     (const cx<T> *)twN);
  return hipGetLastError();
}

hipError_t row_c2r(Dt dtc, Dt dto, int M, const void *in, long in_pitch, long in_plane, int ncols_in,
                   int rows, int nplanes, void *out, long out_pitch, long out_plane, hipStream_t s) {
#define CASE(m)                                                                                        \
  case m:                                                                                              \
    if (dtc == F64 && dto == F64)                                                                      \
      return launch_c2r<double, double, m>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,       \
                                           out_pitch, out_plane, s, dtc);                              \
    if (dtc == F64 && dto == F32)                                                                      \
      return launch_c2r<double, float, m>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,        \
                                          out_pitch, out_plane, s, dtc);                               \
    if (dtc == F32 && dto == F32)                                                                      \
      return launch_c2r<float, float, m>(in, in_pitch, in_plane, ncols_in, rows, nplanes, out,         \
                                         out_pitch, out_plane, s, dtc);                                \
    return hipErrorInvalidValue;
  switch (M) {
    LP_ROW_LENGTHS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

template <typename T, int NYI, int NYO>
static hipError_t launch_col(int padmode, int fwd, int inv, const void *in, long in_pitch, long in_plane,
                             int ncols, int nplanes, int nout, const ColOut *outs, long out_pitch,
                             long out_plane, const SpecTables &tab, hipStream_t s, Dt dt) {
  const void *twI = tw_ptr(NYI, dt), *twO = tw_ptr(NYO, dt);
  if (!twI || !twO) return hipErrorInvalidValue;
  constexpr int NMAX = NYI > NYO ? NYI : NYO;
  constexpr int TX = col_tx(NMAX, (int)sizeof(cx<T>));
  ColParams P;
  P.nout = nout;
  for (int j = 0; j < 3; ++j) P.outs[j] = j < nout ? outs[j] : ColOut{nullptr, 0};
  dim3 grid((unsigned)((ncols + TX - 1) / TX), (unsigned)nplanes);
  hipLaunchKernelGGL((k_col_op<T, NYI, NYO>), grid, dim3(kBlock), 0, s, (const cx<T> *)in, in_pitch, in_plane,
                     ncols, nplanes, padmode, fwd, inv, P, out_pitch, out_plane, tab.kx, tab.ky, tab.G1,
                     tab.G2, tab.lh, tab.ny, (const cx<T> *)twI, (const cx<T> *)twO);
  return hipGetLastError();
}

hipError_t col_op(Dt dt, int nyi, int nyo, int padmode, int fwd, int inv, const void *in, long in_pitch,
                  long in_plane, int ncols, int nplanes, int nout, const ColOut *outs, long out_pitch,
                  long out_plane, const SpecTables &tab, hipStream_t s) {
  if (nout < 1 || nout > 3) return hipErrorInvalidValue;
#define SAME(n)                                                                                         \
  if (nyi == n && nyo == n)                                                                             \
    return dt == F64 ? launch_col<double, n, n>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,       \
                                                nplanes, nout, outs, out_pitch, out_plane, tab, s, dt) \
                     : launch_col<float, n, n>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,        \
                                               nplanes, nout, outs, out_pitch, out_plane, tab, s, dt);
#define PAIR(a, b)                                                                                      \
  if (nyi == a && nyo == b)                                                                             \
    return dt == F64 ? launch_col<double, a, b>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,       \
                                                nplanes, nout, outs, out_pitch, out_plane, tab, s, dt) \
                     : launch_col<float, a, b>(padmode, fwd, inv, in, in_pitch, in_plane, ncols,        \
                                               nplanes, nout, outs, out_pitch, out_plane, tab, s, dt);
  SAME(32) SAME(48) SAME(64) SAME(96) SAME(128) SAME(144) SAME(192) SAME(256) SAME(288) SAME(384)
  PAIR(32, 48) PAIR(48, 32) PAIR(64, 96) PAIR(96, 64) PAIR(96, 144) PAIR(144, 96)
  PAIR(128, 192) PAIR(192, 128) PAIR(192, 288) PAIR(288, 192) PAIR(256, 384) PAIR(384, 256)
#undef SAME
#undef PAIR
  return hipErrorInvalidValue;
}

}  // namespace lp
