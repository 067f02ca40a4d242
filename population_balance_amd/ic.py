"""Seeded synthetic initial condition (host side, numpy, binary64).

The reference starts from a restart file or a log profile plus ran3 noise
(initial.f90:74-140, ic.f90:47-125).  This generator instead produces a discretely
divergence-free turbulent field on the staggered grid, written once and shared by
the GPU and CPU runs:

  * a random vector potential A with complex Gaussian coefficients of magnitude
    |A(k)| ~ k^(-17/6) for 1 <= |k| <= min(nx,ny)/3 (so |u(k)| ~ k^(-11/6), shell
    energy ~ k^(-5/3)); |k| is measured in units of 2*pi/L_x horizontally and
    pi/L_z vertically;
  * A_x, A_y on w-nodes with a sine basis in z (zero at the wall and the lid),
    A_z on uvp-nodes with a cosine basis;
  * u = curl A with spectral x-y derivatives and the staggered z difference, so
    d/dx u + d/dy v + (w(j+1)-w(j))/dz vanishes to rounding, and w(1) = w(nz) = 0;
  * fluctuation rescaled to u_rms = 1 (rms over the three components), then the log
    law U(z) = ln(z/z0)/kappa added to u at z = (jz-1/2) dz; u, v(nz) = u, v(nz-1).

Random numbers: SplitMix64 (Steele, Lea & Flood 2014) with Box-Muller normals.
"""
import numpy as np

DEFAULT_SEED = 20261202
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n consecutive outputs of SplitMix64 started from `seed`, skipping `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, offset=0):
    """doubles in [0,1) from the top 53 bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normals(seed, n, offset=0):
    """n standard normals by Box-Muller on consecutive uniform pairs."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m, offset)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    th = 2.0 * np.pi * u2
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out[:n]


def _zsynth(B, coef):
    """sum_m B[j, m] coef[m, ...] with real B (two real GEMMs)."""
    flat = coef.reshape(coef.shape[0], -1)
    re = B @ np.ascontiguousarray(flat.real)
    im = B @ np.ascontiguousarray(flat.imag)
    return (re + 1j * im).reshape((B.shape[0],) + coef.shape[1:])


def _irfft2_levels(spec, nx, ny, ld):
    """(nl, ny, nx//2+1) complex spectra -> (nl, ny, ld) real planes (padding 0)."""
    phys = np.fft.irfft2(spec, s=(ny, nx), axes=(-2, -1))
    out = np.zeros(spec.shape[:-2] + (ny, ld))
    out[..., :nx] = phys
    return out


def turbulent_ic(grid, seed=DEFAULT_SEED, z0=1e-4, vonk=0.4, u_rms=1.0, log_law=True,
                 kmax=None):
    """Return u, v, w as (nz, ny, ld) binary64 arrays (C view of Fortran (ld,ny,nz))."""
    nx, ny, nz, ld = grid.nx, grid.ny, grid.nz, grid.ld
    Lx, Ly, Lz = grid.L_x, grid.L_y, grid.L_z
    dz = grid.dz
    lh = nx // 2 + 1
    if kmax is None:
        kmax = min(nx, ny) / 3.0
    mx = np.arange(lh).astype(np.float64)
    my = np.fft.fftfreq(ny, 1.0 / ny)
    MY, MX = np.meshgrid(my, mx, indexing="ij")          # (ny, lh)
    mzmax = int(np.floor(kmax))
    mzmax = max(1, min(mzmax, (nz - 1) // 3))
    mz = np.arange(0, mzmax + 1).astype(np.float64)      # 0 only used by A_z
    # |k| in units of the fundamental of each direction (L_x = 2 pi -> integer kx)
    kh2 = (MX * (2 * np.pi / Lx)) ** 2 + (MY * (2 * np.pi / Ly)) ** 2
    kz = mz * np.pi / Lz
    K = np.sqrt(kh2[None, :, :] + kz[:, None, None] ** 2)          # (nmz, ny, lh)
    band = (K >= 1.0) & (K <= kmax) & (np.abs(MX)[None] <= kmax) & (np.abs(MY)[None] <= kmax)
    amp = np.where(band, np.maximum(K, 1e-30) ** (-17.0 / 6.0), 0.0)
    nm = amp.size
    coef = []
    for comp in range(3):
        g = normals(seed, 2 * nm, offset=comp * 2 * nm).reshape(2, *amp.shape)
        coef.append((g[0] + 1j * g[1]) * amp)
    Ax_c, Ay_c, Az_c = coef
    Ax_c[0] = 0.0                                        # sine basis starts at mz = 1
    Ay_c[0] = 0.0
    # z bases: w-nodes z_w = (j-1) dz, uvp-nodes z_u = (j-1/2) dz, j = 1..nz
    zw = np.arange(nz) * dz
    zu = (np.arange(nz) + 0.5) * dz
    Sw = np.sin(np.outer(zw, kz))                        # (nz, nmz)
    Cu = np.cos(np.outer(zu, kz))
    Ax = _zsynth(Sw, Ax_c)                                # spectra on w-nodes (nz, ny, lh)
    Ay = _zsynth(Sw, Ay_c)
    Az = _zsynth(Cu, Az_c)                                # on uvp-nodes
    ikx = 1j * MX * (2 * np.pi / Lx)
    iky = 1j * MY * (2 * np.pi / Ly)
    U = np.zeros((nz, ny, lh), complex)
    V = np.zeros((nz, ny, lh), complex)
    W = np.zeros((nz, ny, lh), complex)
    # uvp level j (0-based k) sits between w-nodes k and k+1
    U[:-1] = iky * Az[:-1] - (Ay[1:] - Ay[:-1]) / dz
    V[:-1] = (Ax[1:] - Ax[:-1]) / dz - ikx * Az[:-1]
    W[:] = ikx * Ay - iky * Ax
    u = _irfft2_levels(U, nx, ny, ld)
    v = _irfft2_levels(V, nx, ny, ld)
    w = _irfft2_levels(W, nx, ny, ld)
    w[0] = 0.0
    w[-1] = 0.0
    inner = slice(0, nz - 1)
    rms = np.sqrt((np.mean(u[inner, :, :nx] ** 2) + np.mean(v[inner, :, :nx] ** 2) +
                   np.mean(w[inner, :, :nx] ** 2)) / 3.0)
    s = u_rms / rms if rms > 0 else 0.0
    u *= s
    v *= s
    w *= s
    if log_law:
        prof = np.log(zu / z0) / vonk
        u[:, :, :nx] += prof[:, None, None]
    u[-1] = u[-2]
    v[-1] = v[-2]
    return u, v, w


def discrete_divergence(grid, u, v, w):
    """d/dx u + d/dy v + (w(jz+1)-w(jz))/dz on uvp levels 1..nz-1 (spectral x-y)."""
    nx, ny = grid.nx, grid.ny
    kx = np.fft.rfftfreq(nx, 1.0 / nx) * 2 * np.pi / grid.L_x
    ky = np.fft.fftfreq(ny, 1.0 / ny) * 2 * np.pi / grid.L_y
    KY, KX = np.meshgrid(ky, kx, indexing="ij")
    KX[:, -1] = 0.0
    KY[ny // 2, :] = 0.0
    U = np.fft.rfft2(u[:-1, :, :nx], axes=(-2, -1))
    V = np.fft.rfft2(v[:-1, :, :nx], axes=(-2, -1))
    d = np.fft.irfft2(1j * KX * U + 1j * KY * V, s=(ny, nx), axes=(-2, -1))
    return d + (w[1:, :, :nx] - w[:-1, :, :nx]) / grid.dz
