// lp_convec.hip -- the 3/2-grid part of convec.f90 as one row pass.
//
// convec.f90 pads u, v, w and the vorticity to the 3/2 grid (dealias1), forms the cross
// products there (:138-158, 180-221, 241-279) and truncates them back (dealias2).  The
// y-padding and y-truncation stay column passes (col_op); everything along x is fused
// here: one block takes one y row of the 3/2 grid and L consecutive product levels,
// back-transforms the rows of the six fields it needs along x (merge + four-step FFT,
// exactly as row_c2r), forms the products in LDS and transforms them forward along x
// (four-step FFT + split, exactly as row_r2c), writing only the kx < nx/2 columns that
// dealias2 keeps.  The products couple neighbouring levels (w, w_x and w_y are averaged
// onto the uv-nodes, u and v onto the w-nodes), so a block transforms
//   u1, u2      at levels k0-1 .. k0+L-1,
//   w3          at levels k0   .. k0+L-1,
//   u3, w1, w2  at levels k0   .. k0+L,
// 6L+5 rows for L levels of products.
#include "lp_convec.hpp"
#include "lp_fft4.hpp"

namespace lp {

// lanes per transform: 16 for the long transforms (Q = 12 for M = 192, 9 for M = 144),
// which keeps the per-lane values of a transform within the register budget
__host__ __device__ constexpr int cv_lanes(int M) { return (M == 192 || M == 144) ? 16 : row_lanes(M); }

template <typename T, int M> struct CvCfg {
  static constexpr int P = cv_lanes(M);
  static constexpr int L = sizeof(T) == 4 ? 4 : 2;  // product levels per block
  static constexpr int NU = L + 1;                  // levels held of u1, u2, u3, w1, w2
  static constexpr int NS = 6 * L + 5;              // x back-transforms per block
  static constexpr int NF = 3 * L;                  // x forward transforms per block
  static constexpr int NTHR = (NS * P + 63) / 64 * 64;
  using FI = Fft4<T, M, P, +1>;
  using FO = Fft4<T, M, P, -1>;
  static constexpr int SSI = FI::SCR | 1;  // unpadded (odd): room for the twiddles
  static constexpr int SSO = slot_stride(FO::SCR > M ? FO::SCR : M);
  static constexpr int RL = 2 * M;                    // physical row length
  static constexpr int NPT = (L * RL + NTHR - 1) / NTHR;  // product points per thread
  static constexpr size_t mx(size_t a, size_t b) { return a > b ? a : b; }
  static constexpr size_t work() {
    return (mx(mx((size_t)NS * SSI * sizeof(cx<T>), (size_t)NS * RL * sizeof(T)),
               mx((size_t)NF * RL * sizeof(T), (size_t)NF * SSO * sizeof(cx<T>))) + 15) / 16 * 16;
  }
  // the transforms' area, then the twiddles twM[0..M), twN[0..M]
  static constexpr size_t lds() { return work() + (2 * M + 1) * sizeof(cx<T>); }
};

template <typename T, int M>
__global__ __launch_bounds__((CvCfg<T, M>::NTHR)) void k_convec_rows(const cx<T> *__restrict__ B, cx<T> *__restrict__ C,
                                                                    long pitch, int ny2, int nz, int ncols, T cb,
                                                                    const cx<T> *__restrict__ twM,
                                                                    const cx<T> *__restrict__ twN) {
  using K = CvCfg<T, M>;
  constexpr int P = K::P, L = K::L, NU = K::NU, NS = K::NS, NF = K::NF, RL = K::RL;
  constexpr int Q = K::FI::Q;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T> *scr = reinterpret_cast<cx<T> *>(smem);  // transform scratch
  T *vals = reinterpret_cast<T *>(smem);         // [NS][RL] rows, then [NF][RL] products
  cx<T> *tws = reinterpret_cast<cx<T> *>(smem + K::work());
  for (int i = threadIdx.x; i <= 2 * M; i += K::NTHR) tws[i] = i < M ? twM[i] : twN[i - M];
  const int y = blockIdx.x % ny2;
  const int k0 = (blockIdx.x / ny2) * L;
  const int slot = threadIdx.x / P, t = threadIdx.x % P;
  const long fplane = (long)nz * ny2 * pitch;  // one field of B or C

  // ---- x back-transforms of the rows this block needs ----
  int f, lev;
  if (slot < NU) {
    f = 0; lev = k0 - 1 + slot;
  } else if (slot < 2 * NU) {
    f = 1; lev = k0 - 1 + (slot - NU);
  } else if (slot < 2 * NU + L) {
    f = 5; lev = k0 + (slot - 2 * NU);
  } else {
    const int r = slot - 2 * NU - L;
    f = 2 + r / NU; lev = k0 + r % NU;
  }
  const bool inv = slot < NS;
  const bool act = inv && lev >= 0 && lev < nz;
  const cx<T> *src = B + (act ? f * fplane + ((long)lev * ny2 + y) * pitch : 0);
  auto ldx = [&](int kk) -> cx<T> {
    cx<T> x = (act && kk < ncols) ? src[kk] : mk<T>(0, 0);
    if (kk == 0 || kk == M) x.y = 0;
    return x;
  };
  {
    // merge (as row_c2r): Y[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k])
    cx<T> xa_[Q], xb_[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int kk = t + P * q;
      xa_[q] = ldx(kk);
      xb_[q] = ldx(M - kk);
    }
    __syncthreads();  // the twiddles
    cx<T> v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int kk = t + P * q;
      cx<T> xa = xa_[q], xb = conjg(xb_[q]);
      cx<T> e = xa + xb, o = xa - xb;
      cx<T> wo = cmul(conjg(tws[M + kk]), o);
      v[q] = mk<T>(e.x - wo.y, e.y + wo.x);
    }
    cx<T> *sc = scr + (inv ? slot : 0) * K::SSI;
    if (inv) K::FI::stage1(v, t, sc, tws);
    slot_sync<P>();
    cx<T> a[K::FI::G][P];
    bool ok[K::FI::G];
#pragma unroll
    for (int j = 0; j < K::FI::G; ++j) ok[j] = inv && K::FI::stage2(j, t, sc, a[j]);
    __syncthreads();  // every transform has read its scratch: the LDS now holds rows
#pragma unroll
    for (int j = 0; j < K::FI::G; ++j)
      if (ok[j]) {
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
          const int n = t + P * j + Q * k1;
          vals[slot * RL + 2 * n] = a[j][k1].x;
          vals[slot * RL + 2 * n + 1] = a[j][k1].y;
        }
      }
  }
  __syncthreads();

  // ---- products (convec.f90:138-158, 180-221, 241-279), normalisation folded in ----
  T pr[K::NPT][3];
  const T h = T(0.5);
#pragma unroll
  for (int m = 0; m < K::NPT; ++m) {
    const int p = threadIdx.x + K::NTHR * m;
    const int i = p / RL, x = p % RL;
    const int k = k0 + i;
    T cxv = T(0), cyv = T(0), czv = T(0);
    if (p < L * RL && k < nz - 1) {
      auto V = [&](int s) -> T { return vals[s * RL + x]; };
      const T u1 = V(i + 1), u2 = V(NU + i + 1), w3 = V(2 * NU + i);
      const T u3a = V(2 * NU + L + i), u3b = V(2 * NU + L + i + 1);
      const T w1a = V(3 * NU + L + i), w1b = V(3 * NU + L + i + 1);
      const T w2a = V(4 * NU + L + i), w2b = V(4 * NU + L + i + 1);
      if (k == 0) {
        cxv = cb * (u2 * (-w3) + h * u3b * (w2b));
        cyv = cb * (u1 * (w3) + h * u3b * (-w1b));
      } else {
        const T u1m = V(i), u2m = V(NU + i);
        cxv = cb * (u2 * (-w3) + h * (u3b * (w2b) + u3a * (w2a)));
        cyv = cb * (u1 * (w3) + h * (u3b * (-w1b) + u3a * (-w1a)));
        czv = cb * h * ((u1 + u1m) * (-w2a) + (u2 + u2m) * (w1a));
      }
    }
    pr[m][0] = cxv;
    pr[m][1] = cyv;
    pr[m][2] = czv;
  }
  __syncthreads();  // rows consumed: products go over them, component-major [c][i][x]
#pragma unroll
  for (int m = 0; m < K::NPT; ++m) {
    const int p = threadIdx.x + K::NTHR * m;
    if (p < L * RL) {
      const int i = p / RL, x = p % RL;
#pragma unroll
      for (int c = 0; c < 3; ++c) vals[(c * L + i) * RL + x] = pr[m][c];
    }
  }
  __syncthreads();

  // ---- x forward transforms of the products (as row_r2c), kx < ncols written ----
  const bool fwd = slot < NF;
  const int c = fwd ? slot / L : 0, i = fwd ? slot % L : 0;
  const int k = k0 + i;
  using v2 = typename V2<T>::t;
  cx<T> z[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const v2 a = fwd ? *reinterpret_cast<const v2 *>(vals + slot * RL + 2 * (t + P * q)) : v2{0, 0};
    z[q] = mk<T>(a.x, a.y);
  }
  __syncthreads();  // products read: the scratch of the forward transforms goes over them
  cx<T> *sc = scr + (fwd ? slot : 0) * K::SSO;
  if (fwd) K::FO::stage1(z, t, sc, tws);
  slot_sync<P>();
  cx<T> a[K::FO::G][P];
  bool ok[K::FO::G];
#pragma unroll
  for (int j = 0; j < K::FO::G; ++j) ok[j] = fwd && K::FO::stage2(j, t, sc, a[j]);
  slot_sync<P>();
#pragma unroll
  for (int j = 0; j < K::FO::G; ++j)
    if (ok[j]) {
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) sc[t + P * j + Q * k1] = a[j][k1];
    }
  slot_sync<P>();
  if (fwd && k < nz - 1) {
    // split: X[k] = (Z[k] + conj Z[M-k])/2 - i/2 w^k (Z[k] - conj Z[M-k]), w = exp(-2 pi i/2M)
    cx<T> *dst = C + c * fplane + ((long)k * ny2 + y) * pitch;
    for (int kk = t; kk < ncols; kk += P) {
      cx<T> za = sc[kk % M], zb = conjg(sc[(M - kk) % M]);
      cx<T> e = za + zb, o = za - zb;
      cx<T> wo = cmul(tws[M + kk], o);
      dst[kk] = mk<T>(T(0.5) * (e.x + wo.y), T(0.5) * (e.y - wo.x));
    }
  }
}

template <typename T, int M>
static hipError_t launch_convec_rows(const void *B, void *C, long pitch, int ny2, int nz, int ncols, double cb,
                                     hipStream_t s, Dt dt) {
  using K = CvCfg<T, M>;
  const void *twM = tw_ptr(M, dt), *twN = tw_ptr(2 * M, dt);
  if (!twM || !twN || ncols < 1 || ncols > M || pitch < ncols || ny2 < 1 || nz < 2) return hipErrorInvalidValue;
  const size_t lds = K::lds();
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_convec_rows<T, M>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const long chunks = (nz - 1 + K::L - 1) / K::L;
  hipLaunchKernelGGL((k_convec_rows<T, M>), dim3((unsigned)(chunks * ny2)), dim3(K::NTHR), lds, s,
                     (const cx<T> *)B, (cx<T> *)C, pitch, ny2, nz, ncols, (T)cb, (const cx<T> *)twM,
                     (const cx<T> *)twN);
  return hipGetLastError();
}

#define LP_BIG_ROWS(X) X(24) X(48) X(72) X(96) X(144) X(192)

hipError_t convec_rows(Dt dt, int M, const void *B, void *C, long pitch, int ny2, int nz, int ncols, double cb,
                       hipStream_t s) {
#define CASE(m)                                                                                      \
  case m:                                                                                            \
    return dt == F64 ? launch_convec_rows<double, m>(B, C, pitch, ny2, nz, ncols, cb, s, dt)         \
                     : launch_convec_rows<float, m>(B, C, pitch, ny2, nz, ncols, cb, s, dt);
  switch (M) {
    LP_BIG_ROWS(CASE)
    default: return hipErrorInvalidValue;
  }
#undef CASE
}

}  // namespace lp
