// lp_lag.hip -- the Lagrangian scale-dependent model's 42 test filters fused with the
// products that feed them and the contractions that consume them
// (lagrange_Sdep.f90:129-434 with interpolag_Sdep.f90:62-365).
//
//   k_lag_r2c    forms field q of a row from u, v, w, S_ij, |S| in registers and does the
//                x transform (four-step FFT + real split), writing only the kx columns
//                below the filters' cutoff (the 21 inputs are never stored as volumes);
//   col_op       (lp_fft.hip) filters all 21 fields along y at both widths in one pass;
//   k_lag_c2r_contract  transforms the 21 fields of one width back along x for one
//                (y, z) row in LDS and writes their two Germano contractions per point;
//   k_lag_relax  interpolates the old averages at the backtraced point and relaxes them
//                with the four contractions, one thread per point, marching a tile
//                through a run of levels (other workgroups relax the padding columns).
#include "lp_fft4.hpp"
#include "lp_lag.hpp"

namespace lp {

template <typename T> struct LagIn {
  const double *u, *v, *w;  // the velocities are binary64 in both precisions
  const T *S[6];
  const T *magS;
};
struct F4c {
  const double *p[4];
};
struct F4 {
  double *p[4];
};

// lagrange_Sdep.f90:129-155,192-242: the 21 filter inputs.  Order: 0-2 u, v, w on
// w-nodes, 3-8 the products 11,12,13,22,23,33, 9-14 S_ij, 15-20 |S| S_ij.  Each helper
// returns the values at the two points i, i+1 (i even) of level k.
template <typename T> struct LagPair {
  using v2 = typename V2<T>::t;
  using d2 = typename V2<double>::t;
  const LagIn<T> &in;
  const long np;
  const int k;
  double cu, cv;  // per-level shifts of u, v on the w-node (zero in binary64)
  __device__ __forceinline__ static v2 ld2(const T *f, long o) { return *reinterpret_cast<const v2 *>(f + o); }
  __device__ __forceinline__ static d2 ldd(const double *f, long o) {
    return *reinterpret_cast<const d2 *>(f + o);
  }
  __device__ __forceinline__ static v2 mul(v2 a, v2 b) { return v2{a.x * b.x, a.y * b.y}; }
  // branch-free (loads from clamped offsets, then a select) so that the loads of a row
  // are issued together
  __device__ __forceinline__ d2 hv(const double *f, long i) const {  // u, v on the w-node (jz = 1 stays)
    const d2 a = ldd(f, i);
    const d2 b = ldd(f, i - (k > 0 ? np : 0));
    const d2 h{0.5 * (a.x + b.x), 0.5 * (a.y + b.y)};
    return k == 0 ? a : h;
  }
  __device__ __forceinline__ v2 hs(const double *f, double c, long i) const {  // hv(f, i) - c, rounded to T
    const d2 a = hv(f, i);
    return v2{(T)(a.x - c), (T)(a.y - c)};
  }
  __device__ __forceinline__ v2 wn(long i) const {  // w; at jz = 1 a quarter of w(2)
    const d2 a = ldd(in.w, i + (k == 0 ? np : 0));
    const d2 q{0.25 * a.x, 0.25 * a.y};
    return k == 0 ? v2{(T)q.x, (T)q.y} : v2{(T)a.x, (T)a.y};
  }
};

// z[m] = field values at points 2(t + P m), 2(t + P m) + 1 of the row starting at base;
// the loads of all m are independent, so they are issued together
template <typename T, int Q, int P, typename Fn>
__device__ __forceinline__ void row_pairs(Fn f, cx<T> (&z)[Q], long base, int t) {
#pragma unroll
  for (int m = 0; m < Q; ++m) {
    const auto a = f(base + 2 * (t + P * m));
    z[m] = mk<T>(a.x, a.y);
  }
}

template <typename T, int Q, int P>
__device__ __forceinline__ void lag_row(int q, const LagIn<T> &in, const Geo &g, int k, long base, int t,
                                        cx<T> (&z)[Q]) {
  // Binary32: u and v enter shifted by their w-node value at the level's first point.  Only
  // L_ij = bar(u_i u_j) - bar(u_i) bar(u_j) reads fields 0-8, and it is unchanged by a
  // constant shift (both test filters pass kx = ky = 0 with gain 1); without the mean flow
  // in the products that difference no longer cancels |U|^2 / |u'|^2 of binary32 rounding.
  // Binary64 keeps the reference's products (shift 0, so every value is unchanged).
  LagPair<T> L{in, g.np, k, 0.0, 0.0};
  if constexpr (sizeof(T) == 4) {
    if (q <= 7 && q != 2) {  // the fields that read u or v
      L.cu = L.hv(in.u, (long)k * g.np).x;
      L.cv = L.hv(in.v, (long)k * g.np).x;
    }
  }
  using LP = LagPair<T>;
  switch (q) {  // q is uniform over the wavefront
    case 0: row_pairs<T, Q, P>([&](long i) { return L.hs(in.u, L.cu, i); }, z, base, t); break;
    case 1: row_pairs<T, Q, P>([&](long i) { return L.hs(in.v, L.cv, i); }, z, base, t); break;
    case 2: row_pairs<T, Q, P>([&](long i) { return L.wn(i); }, z, base, t); break;
    case 3: row_pairs<T, Q, P>([&](long i) { auto a = L.hs(in.u, L.cu, i); return LP::mul(a, a); }, z, base, t); break;
    case 4: row_pairs<T, Q, P>([&](long i) { return LP::mul(L.hs(in.u, L.cu, i), L.hs(in.v, L.cv, i)); }, z, base, t); break;
    case 5: row_pairs<T, Q, P>([&](long i) { return LP::mul(L.hs(in.u, L.cu, i), L.wn(i)); }, z, base, t); break;
    case 6: row_pairs<T, Q, P>([&](long i) { auto a = L.hs(in.v, L.cv, i); return LP::mul(a, a); }, z, base, t); break;
    case 7: row_pairs<T, Q, P>([&](long i) { return LP::mul(L.hs(in.v, L.cv, i), L.wn(i)); }, z, base, t); break;
    case 8: row_pairs<T, Q, P>([&](long i) { auto a = L.wn(i); return LP::mul(a, a); }, z, base, t); break;
    default:
      if (q < 15) {
        const T *S = in.S[q - 9];
        row_pairs<T, Q, P>([&](long i) { return LP::ld2(S, i); }, z, base, t);
      } else {
        const T *S = in.S[q - 15];
        row_pairs<T, Q, P>([&](long i) { return LP::mul(LP::ld2(in.magS, i), LP::ld2(S, i)); }, z, base, t);
      }
  }
}

// One wavefront transforms one field for 64/P rows; the four wavefronts of a block take
// four fields at a time, six rounds cover the 21.
template <typename T, int M>
__global__ __launch_bounds__(256) void k_lag_r2c(LagIn<T> in, Geo g, cx<T> *__restrict__ spec, int nc, long pitch,
                                                 const cx<T> *__restrict__ twM, const cx<T> *__restrict__ twN) {
  constexpr int P = row_lanes(M);
  using F = Fft4<T, M, P, -1>;
  constexpr int Q = F::Q;
  constexpr int RB = 64 / P;
  constexpr int SS = slot_stride(F::SCR > M ? F::SCR : M);
  __shared__ cx<T> lds[4 * RB * SS];
  __shared__ cx<T> tws[2 * M + 1];  // twM[0..M), twN[0..M]
  for (int i = threadIdx.x; i <= 2 * M; i += 256) tws[i] = i < M ? twM[i] : twN[i - M];
  __syncthreads();
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l / P, t = l % P;
  const long nrows = (long)g.ny * g.nz;
  const long row = (long)blockIdx.x * RB + r;
  const bool valid = row < nrows;
  const int k = valid ? (int)(row / g.ny) : 0, y = valid ? (int)(row % g.ny) : 0;
  const long base = (long)k * g.np + (long)y * g.ld;
  cx<T> *sc = lds + (wv * RB + r) * SS;
  for (int it = 0; it < 6; ++it) {
    const int q = __builtin_amdgcn_readfirstlane(4 * it + wv);
    const bool act = valid && q < 21;
    cx<T> z[Q];
    if (act) {
      lag_row<T, Q, P>(q, in, g, k, base, t, z);
    } else {
#pragma unroll
      for (int m = 0; m < Q; ++m) z[m] = mk<T>(0, 0);
    }
    F::stage1(z, t, sc, tws);
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
    if (act) {  // split (as k_row_r2c), leading nc columns
      cx<T> *dst = spec + (((long)q * g.nz + k) * g.ny + y) * pitch;
      for (int kk = t; kk < nc; kk += P) {
        cx<T> za = sc[kk % M], zb = conjg(sc[(M - kk) % M]);
        cx<T> e = za + zb, o = za - zb;
        cx<T> wo = cmul(tws[M + kk], o);
        dst[kk] = mk<T>(T(0.5) * (e.x + wo.y), T(0.5) * (e.y - wo.x));
      }
    }
    slot_sync<P>();
  }
}

// lagrange_Sdep.f90:257-326: x back-transform of the 21 fields of one width for one
// (y, z) row and their Germano contractions.  blockIdx.x = row, blockIdx.y = width
// (0: 2 delta -> L.M, M.M; 1: 4 delta -> Q.N, N.N).  21 transforms of P lanes each in
// LDS slots packed without padding (fewer bytes, more workgroups per CU); each transform
// writes its physical row into its own slot once it has read it (a slot is long enough:
// 2 (M + Q) >= ld), so only a wavefront barrier separates the two; the twiddles sit behind
// the slots, so that the transforms issue no global loads.  c0/c1: the two contractions
// of each width.
template <typename T, int M>
__global__ __launch_bounds__(256) void k_lag_c2r_contract(const cx<T> *__restrict__ bar, const cx<T> *__restrict__ hat,
                                                          int n1, int n2, long p1, long p2, T *__restrict__ lm,
                                                          T *__restrict__ mm, T *__restrict__ qn,
                                                          T *__restrict__ nn, Geo g, double cst, double tf1_2,
                                                          double tf2_2, const cx<T> *__restrict__ twM,
                                                          const cx<T> *__restrict__ twN, int tw_off) {
  constexpr int PL = row_lanes(M);
  using F = Fft4<T, M, PL, +1>;
  constexpr int Q = F::Q;
  constexpr int SS = F::SCR;
  using v2 = typename V2<T>::t;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T> *scr = reinterpret_cast<cx<T> *>(smem);  // [21][SS] during the transforms
  T *vals = reinterpret_cast<T *>(smem);         // [21][rs]: each slot's row in the slot
  const int ldp = g.ld;
  constexpr int rs = 2 * SS;  // >= ld = 2M + 2 (SS = M + Q)
  const int row = blockIdx.x, half = blockIdx.y;
  const int k = row / g.ny, y = row % g.ny;
  const int slot = threadIdx.x / PL, t = threadIdx.x % PL;
  const bool act = slot < 21;
  const int q = act ? slot : 0;
  const cx<T> *spec = half == 0 ? bar : hat;
  const int nin = half == 0 ? n1 : n2;
  const long pitch = half == 0 ? p1 : p2;
  const cx<T> *src = spec + (((long)q * g.nz + k) * g.ny + y) * pitch;
  // twiddles in LDS, behind the transforms' slots (2M values, read by every lane)
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem + tw_off);
  for (int i = threadIdx.x; i < 2 * M; i += blockDim.x) tw[i] = i < M ? twM[i] : twN[i - M];
  // merge (as k_row_c2r): Y[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k]).  The
  // loads of a q2 whose columns all lie at or past nin for every lane are skipped by a
  // uniform branch.
  cx<T> xa_[Q], xb_[Q];
#pragma unroll
  for (int q2 = 0; q2 < Q; ++q2) {
    const int kk = t + PL * q2;
    xa_[q2] = mk<T>(0, 0);
    xb_[q2] = mk<T>(0, 0);
    if (PL * q2 < nin) {  // uniform
      if (act && kk < nin) xa_[q2] = src[kk];
    }
    if (M - PL * q2 - (PL - 1) < nin) {  // uniform: some lane has M - kk < nin
      if (act && M - kk < nin) xb_[q2] = src[M - kk];
    }
  }
  __syncthreads();  // the twiddles
  cx<T> v[Q];
#pragma unroll
  for (int q2 = 0; q2 < Q; ++q2) {
    const int kk = t + PL * q2;
    cx<T> xa = xa_[q2], xb = xb_[q2];
    if (kk == 0) { xa.y = 0; xb.y = 0; }  // X[0] and X[M] (M - kk == M only for kk == 0)
    xb = conjg(xb);
    cx<T> e = xa + xb, o = xa - xb;
    cx<T> wo = cmul(conjg(tw[M + kk]), o);
    v[q2] = mk<T>(e.x - wo.y, e.y + wo.x);
  }
  cx<T> *sc = scr + (act ? slot : 0) * SS;
  if (act) F::stage1(v, t, sc, tw);
  slot_sync<PL>();
  cx<T> a[F::G][PL];
  bool ok[F::G];
#pragma unroll
  for (int j = 0; j < F::G; ++j) ok[j] = act && F::stage2(j, t, sc, a[j]);
  slot_sync<PL>();  // the slot's transform has read it: the slot now holds its row
#pragma unroll
  for (int j = 0; j < F::G; ++j)
    if (ok[j]) {
#pragma unroll
      for (int k1 = 0; k1 < PL; ++k1) {
        const int n = t + PL * j + Q * k1;
        *reinterpret_cast<v2 *>(vals + q * rs + 2 * n) = v2{a[j][k1].x, a[j][k1].y};
      }
    }
  if (act)
    for (int x = 2 * M + t; x < ldp; x += PL) vals[q * rs + x] = T(0);
  __syncthreads();
  T *o0 = half == 0 ? lm : qn, *o1 = half == 0 ? mm : nn;
  const double tf = half == 0 ? tf1_2 : tf2_2;
  const long base = (long)k * g.np + (long)y * g.ld;
  for (int x = threadIdx.x; x < ldp; x += blockDim.x) {
    double c0, c1;
    contract_point<T>([&](int qq) -> T { return vals[qq * rs + x]; }, cst, tf, c0, c1);
    o0[base + x] = (T)c0;  // c0, c1 are values of T: exact
    o1[base + x] = (T)c1;
  }
}

// k_lag_relax works on tiles of kRx x kRy points and marches each tile through a run of
// levels (2.5D).  The old averages it gathers (4 fields, the point's x/y neighbours on
// both sides, levels k-1..k+1) sit in a ring of three level slots in LDS, each level
// loaded once per block with coalesced loads: slot (sx, sy) of a level holds the value at
// x = wrap(x0-1+sx), y = wrap(y0-1+sy), so the backtrace corner x + addx of local point
// lx is slot lx+1+addx, also across the periodic seam and in a partial last tile.
// Virtual level m (k-1 <= m <= k+1 while level k is relaxed; m = -1 and m = nz occur)
// lives in ring slot (m+3) % 3 and holds level clamp(m).  While level k is relaxed, the
// loads of level k+2 and of the point values of level k+1 are in flight.
constexpr int kRx = 64, kRy = 4;  // kRx = one wavefront, kRy wavefronts per block
constexpr int kHx = kRx + 2, kHy = kRy + 2;
constexpr int kLagRows = 4 * kHy;  // staged rows of one level: 4 fields x kHy
constexpr int kRz = 16;            // levels per block, at most (the runs are balanced)

// The padding columns x = nx .. ld-1, which interpolag_Sdep.f90 leaves alone: point t of
// the (ld - nx) ny nz padding points is relaxed from its old values (no interpolation).
template <typename T>
__device__ __forceinline__ void relax_pad_point(long t, const T *__restrict__ lm, const T *__restrict__ mm,
                                                const T *__restrict__ qn, const T *__restrict__ nn,
                                                const F4c &Fold, const F4 &Fnew, T *__restrict__ Cs,
                                                T *__restrict__ Beta, const Geo &g, const LagParams &P) {
  const int npad = g.ld - g.nx;
  const int px = g.nx + (int)(t % npad);
  const long row = t / npad;  // k * ny + y
  const int k = (int)(row / g.ny);
  const long i = row * g.ld + px;
  double fi[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) fi[q] = Fold.p[q][i];
  T cs, b;
  relax_point<T>((double)lm[i], (double)mm[i], (double)qn[i], (double)nn[i], fi, px >= g.ld - 2, k == g.nz - 1,
                 P, cs, b);
#pragma unroll
  for (int q = 0; q < 4; ++q) Fnew.p[q][i] = fi[q];
  Cs[i] = cs;
  Beta[i] = b;
}

// lagrange_Sdep.f90:327-434 with interpolag_Sdep.f90:62-365, one thread per point and
// level: the old averages are gathered at the backtraced point (from the LDS ring),
// relaxed with the contractions and written to Fnew.  Blocks are dealt to the XCDs
// round-robin by workgroup id; XCD j takes tiles [j*cg, (j+1)*cg) of every level run, so
// the halo columns and rows a tile shares with its neighbours are read from its L2.
// The first npb workgroups (a multiple of 8, so that the tiles keep their XCDs) relax the
// padding columns instead, 256 points each, beside the tiles: a separate branch, so that
// the tiles' relaxation loop holds no loads of its own (a wait for those would also wait
// for the next level's loads in flight).  The interpolation is interp_point's, term for
// term.
template <typename T, bool UNIT, typename TV = T>
__global__ __launch_bounds__(256) void k_lag_relax(const T *__restrict__ lm, const T *__restrict__ mm,
                                                   const T *__restrict__ qn, const T *__restrict__ nn,
                                                   F4c Fold, F4 Fnew, const TV *__restrict__ ul,
                                                   const TV *__restrict__ vl, const TV *__restrict__ wl,
                                                   T *__restrict__ Cs, T *__restrict__ Beta, Geo g, LagParams P,
                                                   LagInterp I, int ntx, int cg, int nkc, int npb) {
  __shared__ double sF[3][kLagRows][kHx];  // [ring slot][field * kHy + sy][sx]
  if ((int)blockIdx.x < npb) {  // uniform over the block
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t < (long)(g.ld - g.nx) * g.ny * g.nz) relax_pad_point<T>(t, lm, mm, qn, nn, Fold, Fnew, Cs, Beta, g, P);
    return;
  }
  const int ntile = ntx * (g.ny / kRy);
  const int wg = (int)blockIdx.x - npb;
  const int j = wg & 7, s = wg >> 3;
  const int kc = s / cg, tt = j * cg + s % cg;
  if (kc >= nkc || tt >= ntile) return;  // uniform over the block
  const int kb = (int)((long)kc * g.nz / nkc), ke = (int)((long)(kc + 1) * g.nz / nkc);
  const int x0 = (tt % ntx) * kRx, y0 = (tt / ntx) * kRy;
  const int lane = threadIdx.x % 64, wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  auto wrapx = [&](int xg) {
    xg = xg < 0 ? xg + g.nx : xg;
    xg = xg >= g.nx ? xg - g.nx : xg;
    return xg >= g.nx ? xg - g.nx : xg;
  };
  // staging of one level: wavefront wv loads rows wv, wv+4, ... (row index uniform over
  // the wavefront), lane l slot l; slots 64, 65 of all rows by threads 0 .. 2 kLagRows - 1.
  // x0 - 1 <= x <= x0 + 64 < 3 nx; the slots past a partial tile only need to stay in
  // bounds (no point reads them).
  auto row_off = [&](int r) {  // r = q*kHy + sy: offset of the row within a level
    int yg = y0 - 1 + r % kHy;
    yg = yg < 0 ? yg + g.ny : (yg >= g.ny ? yg - g.ny : yg);
    return (long)yg * g.ld;
  };
  const int xs = wrapx(x0 - 1 + lane);
  const bool extra = threadIdx.x < 2 * kLagRows;
  const int xr = extra ? (int)(threadIdx.x >> 1) : 0, xsx = 64 + (int)(threadIdx.x & 1);
  // field pointers by selection (an indexed read of the argument array would be a memory
  // load, waited for before the loads it addresses)
  auto fold_q = [&](int q) -> const double * {
    return q == 0 ? Fold.p[0] : (q == 1 ? Fold.p[1] : (q == 2 ? Fold.p[2] : Fold.p[3]));
  };
  constexpr int kPer = kLagRows / 4;
  const double *rowp[kPer];  // the rows this wavefront stages, within level 0
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int r = 4 * it + wv;
    rowp[it] = fold_q(r / kHy) + row_off(r);
  }
  const double *xrow = fold_q(xr / kHy) + row_off(xr) + wrapx(x0 - 1 + xsx);
  auto load_level = [&](int m, double (&h)[kPer], double &hx) {
    const long kg = (long)(m < 0 ? 0 : (m > g.nz - 1 ? g.nz - 1 : m)) * g.np;
#pragma unroll
    for (int it = 0; it < kPer; ++it) h[it] = rowp[it][kg + xs];
    hx = extra ? xrow[kg] : 0.0;
  };
  auto store_level = [&](int m, const double (&h)[kPer], double hx) {
    double(*sl)[kHx] = sF[(m + 3) % 3];
#pragma unroll
    for (int it = 0; it < kPer; ++it) sl[4 * it + wv][lane] = h[it];
    if (extra) sl[xr][xsx] = hx;
  };
  const int lx = threadIdx.x % kRx, ly = threadIdx.x / kRx;
  const int x = x0 + lx, y = y0 + ly;
  const bool pt = x < g.nx;
  const long ij = (long)y * g.ld + (pt ? x : 0);
  // the point values of one level: contractions, lag velocities (u, v at the level, w at
  // the level or, near the wall, at the second w level)
  struct Pt {
    T c[4];
    TV u, v, w;
  };
  auto load_pt = [&](int k, Pt &p) {
    const long i = (long)k * g.np + ij;
    p.c[0] = lm[i]; p.c[1] = mm[i]; p.c[2] = qn[i]; p.c[3] = nn[i];
    p.u = ul[i]; p.v = vl[i];
    p.w = wl[k <= 1 ? (long)g.np + ij : i];
  };
  auto dcs = [&](double v) { return UNIT ? v : v / I.cs; };
  auto relax_store = [&](int k, long i, int xx, const T (&cc)[4], double (&fi)[4]) {
    T cs, b;
    relax_point<T>((double)cc[0], (double)cc[1], (double)cc[2], (double)cc[3], fi, xx >= g.ld - 2,
                   k == g.nz - 1, P, cs, b);
#pragma unroll
    for (int q = 0; q < 4; ++q) Fnew.p[q][i] = fi[q];
    Cs[i] = cs;
    Beta[i] = b;
  };
  // prologue: virtual levels kb-1, kb, kb+1 in the ring; the point values of kb and, for
  // the w-node average of the lag velocities, u and v of level kb-1
  {
    double h[kPer], hx;
    for (int m = kb - 1; m <= kb + 1; ++m) {
      load_level(m, h, hx);
      store_level(m, h, hx);
    }
  }
  Pt cur;
  load_pt(kb, cur);
  TV up = 0, vp = 0;
  if (kb > 0) {
    up = ul[(long)(kb - 1) * g.np + ij];
    vp = vl[(long)(kb - 1) * g.np + ij];
  }
  // the prologue's loads complete here: the loop's prefetches are then the only loads in
  // flight while a level is relaxed, and the counter waits the compiler places in the loop
  // body cover those alone
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int k = kb; k < ke; ++k) {
    const bool more = k + 1 < ke;  // uniform over the block
    double h[kPer], hx = 0.0;
    Pt nxt;
    if (more) {
      load_level(k + 2, h, hx);
      load_pt(k + 1, nxt);
    }
    if (pt) {
      const long i = (long)k * g.np + ij;
      // interpolag_Sdep.f90:204-258 lag velocities on w-nodes (u_lag at jz=1 stays on the uvp node)
      const double ut = dcs((double)cur.u), vt = dcs((double)cur.v);
      double ulag, vlag;
      if (k == 0) {
        ulag = ut;
        vlag = vt;
      } else {
        ulag = 0.5 * (ut + dcs((double)up));
        vlag = 0.5 * (vt + dcs((double)vp));
      }
      const double xp = -ulag * I.sx;  // :263-265
      const double yp = -vlag * I.sy;
      double zp;
      if (k <= 1) {  // :281-291
        const double w2 = dcs((double)cur.w);
        double z2 = -w2 * I.sz;
        z2 = z2 + fmin(z2, 0.0);
        z2 = fmin(1.0, z2);
        z2 = fmax(-1.0, z2);
        zp = z2;
      } else {
        const double wlag = dcs((double)cur.w);
        zp = -wlag * I.sz;
      }
      const int addx = signbit(xp) ? -1 : 1, addy = signbit(yp) ? -1 : 1, addz = signbit(zp) ? -1 : 1;
      const double cxp = fabs(xp), cyp = fabs(yp), czp = fabs(zp);
      const double fx = 1.0 - cxp, fy = 1.0 - cyp, fz = 1.0 - czp;
      // slots of the corners: x, x + addx (periodic); y, y + addy (periodic); level k and
      // virtual level k + addz (holding clamp(k + addz))
      const int sx0 = lx + 1, sxa = lx + 1 + addx, sy0 = ly + 1, sya = ly + 1 + addy;
      double(*s0)[kHx] = sF[k % 3];
      double(*sa)[kHx] = sF[(k + addz + 3) % 3];
      double fi[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r0 = q * kHy + sy0, ra = q * kHy + sya;
        fi[q] = fx * fy * (s0[r0][sx0] * fz + sa[r0][sx0] * czp)
              + fx * cyp * (s0[ra][sx0] * fz + sa[ra][sx0] * czp)
              + cxp * fy * (s0[r0][sxa] * fz + sa[r0][sxa] * czp)
              + cxp * cyp * (s0[ra][sxa] * fz + sa[ra][sxa] * czp);
      }
      relax_store(k, i, x, cur.c, fi);
    }
    if (more) {
      __syncthreads();  // every wavefront is done with the slot of virtual level k-1
      store_level(k + 2, h, hx);
      __syncthreads();
      up = cur.u;
      vp = cur.v;
      cur = nxt;
    }
  }
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
#define LP_SMALL_ROWS(X) X(16) X(32) X(48) X(64) X(96) X(128)

template <typename T, int M>
static hipError_t launch_lag_r2c(const LagInputs &in, Geo g, void *spec, int nc, long pitch, hipStream_t s,
                                 Dt dt) {
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN || nc < 1 || nc > M + 1 || pitch < nc) return hipErrorInvalidValue;
  LagIn<T> a;
  a.u = (const double *)in.u; a.v = (const double *)in.v; a.w = (const double *)in.w;
  for (int q = 0; q < 6; ++q) a.S[q] = (const T *)in.S[q];
  a.magS = (const T *)in.magS;
  constexpr int RB = 64 / row_lanes(M);
  const long rows = (long)g.ny * g.nz;
  hipLaunchKernelGGL((k_lag_r2c<T, M>), dim3((unsigned)((rows + RB - 1) / RB)), dim3(256), 0, s, a, g,
                     (cx<T> *)spec, nc, pitch, (const cx<T> *)twM, (const cx<T> *)twN);
  return hipGetLastError();
}

hipError_t lag_r2c(Dt dt, const LagInputs &in, Geo g, void *spec, int nc, long pitch, hipStream_t s) {
#define CASE(m)                                                                          \
  case m:                                                                                \
    return dt == F64 ? launch_lag_r2c<double, m>(in, g, spec, nc, pitch, s, dt)          \
                     : launch_lag_r2c<float, m>(in, g, spec, nc, pitch, s, dt);
  switch (g.nx / 2) {
    LP_SMALL_ROWS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

template <typename T, int M>
static hipError_t launch_lag_contract(const void *bar, const void *hat, int n1, int n2, long p1, long p2,
                                      void *const C[4], Geo g, const LagParams &P, hipStream_t s, Dt dt) {
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN || n1 < 1 || n2 < 1 || n1 > M + 1 || n2 > M + 1) return hipErrorInvalidValue;
  constexpr int PL = row_lanes(M);
  using F = Fft4<T, M, PL, +1>;
  const size_t lds_fft = (size_t)21 * F::SCR * sizeof(cx<T>);  // the rows live in the slots
  if (2 * F::SCR < g.ld) return hipErrorInvalidValue;
  const size_t lds_a = (lds_fft + 15) / 16 * 16;
  const size_t lds = lds_a + 2 * M * sizeof(cx<T>);
  const int threads = (21 * PL + 63) / 64 * 64;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_lag_c2r_contract<T, M>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_lag_c2r_contract<T, M>), dim3((unsigned)(g.ny * g.nz), 2), dim3(threads), lds, s,
                     (const cx<T> *)bar, (const cx<T> *)hat, n1, n2, p1, p2, (T *)C[0], (T *)C[1], (T *)C[2],
                     (T *)C[3], g, P.cst,
                     P.tf1_2, P.tf2_2, (const cx<T> *)twM, (const cx<T> *)twN, (int)lds_a);
  return hipGetLastError();
}

hipError_t lag_c2r_contract(Dt dt, const void *bar, const void *hat, int n1, int n2, long p1, long p2,
                            void *const C[4], Geo g, const LagParams &P, hipStream_t s) {
#define CASE(m)                                                                                       \
  case m:                                                                                             \
    return dt == F64 ? launch_lag_contract<double, m>(bar, hat, n1, n2, p1, p2, C, g, P, s, dt)       \
                     : launch_lag_contract<float, m>(bar, hat, n1, n2, p1, p2, C, g, P, s, dt);
  switch (g.nx / 2) {
    LP_SMALL_ROWS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

template <typename T, typename TV>
static hipError_t launch_lag_relax(void *const C[4], const double *const Fold[4], double *const Fnew[4],
                                   const void *ul, const void *vl, const void *wl, void *Cs, void *Beta, Geo g,
                                   const LagParams &P, const LagInterp &I, hipStream_t s) {
  F4c fo;
  F4 fn;
  for (int q = 0; q < 4; ++q) { fo.p[q] = Fold[q]; fn.p[q] = Fnew[q]; }
  // the w level read near the wall (index 1) must exist: nz >= 2; relax_point takes the
  // reference's exponents as constants
  if (g.ny % kRy != 0 || g.nz < 2) return hipErrorInvalidValue;
  if (P.powcoeff != -0.125 || P.bexp != 1.0) return hipErrorInvalidValue;
  const int ntx = (g.nx + kRx - 1) / kRx, cg = (ntx * (g.ny / kRy) + 7) / 8;
  const int nkc = (g.nz + kRz - 1) / kRz;  // level runs of nz/nkc levels (rounded)
  const long npts = (long)(g.ld - g.nx) * g.ny * g.nz;
  const int npb = (int)((npts + 255) / 256 + 7) / 8 * 8;
  const dim3 grid((unsigned)(npb + 8L * cg * nkc));
  if (I.cs == 1.0)
    hipLaunchKernelGGL((k_lag_relax<T, true, TV>), grid, dim3(256), 0, s, (const T *)C[0], (const T *)C[1],
                       (const T *)C[2], (const T *)C[3], fo, fn, (const TV *)ul, (const TV *)vl, (const TV *)wl,
                       (T *)Cs, (T *)Beta, g, P, I, ntx, cg, nkc, npb);
  else
    hipLaunchKernelGGL((k_lag_relax<T, false, TV>), grid, dim3(256), 0, s, (const T *)C[0], (const T *)C[1],
                       (const T *)C[2], (const T *)C[3], fo, fn, (const TV *)ul, (const TV *)vl, (const TV *)wl,
                       (T *)Cs, (T *)Beta, g, P, I, ntx, cg, nkc, npb);
  return hipGetLastError();
}

hipError_t lag_relax(Dt dt, void *const C[4], const double *const Fold[4], double *const Fnew[4], const void *ul,
                     const void *vl, const void *wl, void *Cs, void *Beta, Geo g, const LagParams &P,
                     const LagInterp &I, hipStream_t s, bool vel_f64) {
  if (dt == F64) return launch_lag_relax<double, double>(C, Fold, Fnew, ul, vl, wl, Cs, Beta, g, P, I, s);
  return vel_f64 ? launch_lag_relax<float, double>(C, Fold, Fnew, ul, vl, wl, Cs, Beta, g, P, I, s)
                 : launch_lag_relax<float, float>(C, Fold, Fnew, ul, vl, wl, Cs, Beta, g, P, I, s);
}

}  // namespace lp
