"""Pins of the binary64 oracle (CPU).  The reference holds no golden vectors except
the TRI_DEBUG known-answer system (press_stag_array.f90:141-188,245-254), so every
routine is pinned here by a library routine (numpy's pocketfft), a closed form, a
brute-force solve or an invariant, chosen so that a dropped term, a wrong sign or a
wrong index fails."""
import math

import numpy as np
import pytest

from oracle import Oracle
from population_balance_amd import abi, ic

TWO_PI = 2 * math.pi


def _grid_xy(g, big=False):
    nx, ny = (g.nx2, g.ny2) if big else (g.nx, g.ny)
    x = np.arange(nx) * g.L_x / nx
    y = np.arange(ny) * g.L_y / ny
    Y, X = np.meshgrid(y, x, indexing="ij")
    return X, Y


def _plane(g, vals, big=False):
    ld = g.ld_big if big else g.ld
    out = np.zeros(vals.shape[:-1] + (ld,))
    out[..., :vals.shape[-1]] = vals
    return out


@pytest.fixture(scope="module")
def o32():
    cfg = abi.make_config(32, 32, 9)
    return Oracle(cfg)


# ---------------- FFT conventions (fft.f90 / FFTW2 rfftwnd, unnormalised) ----------------
@pytest.mark.parametrize("n", [32, 48, 64, 96, 128, 192, 256, 288, 384])
def test_cfft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out = np.zeros(n, complex)
    o = Oracle(abi.make_config(32, 32, 9))
    o.L.or_cfft1(n, -1, x.ctypes.data, out.ctypes.data)
    assert np.allclose(out, np.fft.fft(x), rtol=0, atol=1e-12 * np.abs(x).sum())
    o.L.or_cfft1(n, +1, x.ctypes.data, out.ctypes.data)
    assert np.allclose(out, np.fft.ifft(x) * n, rtol=0, atol=1e-12 * np.abs(x).sum())


@pytest.mark.parametrize("nx", [32, 64, 128, 256])
def test_rfft2_matches_numpy(nx):
    cfg = abi.make_config(nx, nx, 5)
    o = Oracle(cfg)
    g = cfg.grid
    rng = np.random.default_rng(nx)
    for big in (False, True):
        n = g.nx2 if big else g.nx
        x = rng.standard_normal((n, n))
        F = o.rfft2_fwd(_plane(g, x, big), big)
        assert np.abs(F - np.fft.rfft2(x)).max() < 1e-13 * np.abs(x).sum()
        # c2r of an arbitrary (non-Hermitian) half spectrum: FFTW2 / pocketfft use only the
        # real parts of the kx = 0 and kx = nx/2 terms after the y transform
        S = rng.standard_normal(F.shape) + 1j * rng.standard_normal(F.shape)
        r = o.rfft2_inv(S, big)
        ref = np.fft.irfft2(S, s=(n, n)) * n * n
        assert np.abs(r[:, :n] - ref).max() < 1e-12 * np.abs(ref).max()
        assert np.all(r[:, n:] == 0.0)


def test_wavenumbers_closed_form(o32):
    """fft.f90:51-76: kx = 2pi/L_x (jx-1), ky = 2pi/L_y (wrapped), Nyquist zeroed."""
    cfg = abi.make_config(32, 16, 5, L_x=3.0, L_y=5.0)
    o = Oracle(cfg)
    g = cfg.grid
    kx, ky, k2 = o.init_wavenumber()
    ex = np.tile((TWO_PI / 3.0) * np.arange(g.lh), (g.ny, 1))
    m = np.fft.fftfreq(g.ny, 1 / g.ny)
    ey = np.tile(((TWO_PI / 5.0) * m)[:, None], (1, g.lh))
    for e in (ex, ey):
        e[:, -1] = 0.0
        e[g.ny // 2, :] = 0.0
    assert np.array_equal(kx, ex) and np.array_equal(ky, ey)
    assert np.array_equal(k2, ex ** 2 + ey ** 2)


def test_test_filter_init_cutoff(o32):
    """test_filtermodule.f90:29-57: sharp cutoff at (pi/(alpha sqrt(dx dy)))^2."""
    g = o32.g
    kx, ky, k2 = o32.init_wavenumber()
    for alpha in (2.0, 4.0):
        G = o32.test_filter_init(alpha)
        kc2 = (math.pi / (alpha * math.sqrt(g.dx * g.dy))) ** 2
        # |k| in index units: kc = nx/(2 alpha) for a 2pi box
        assert math.isclose(math.sqrt(kc2), g.nx / (2 * alpha))
        inside = (k2 < kc2)
        inside[:, -1] = False
        inside[g.ny // 2, :] = False
        assert np.array_equal(G != 0, inside)
        assert np.allclose(G[inside], 1.0 / (g.nx * g.ny))


# ---------------- spectral derivatives (filt_da, ddx, ddy, ddxy) ----------------
def test_spectral_derivatives_closed_form(o32):
    g = o32.g
    X, Y = _grid_xy(g)
    f = np.sin(3 * X) * np.cos(5 * Y) + 0.5 * np.cos(7 * X + 2 * Y) + 2.0
    nyq = np.cos(16 * X)                 # kx = nx/2 mode: removed by filt_da
    F = _plane(g, np.stack([f + nyq] * 3))
    ff, dx, dy = o32.filt_da(F)
    ex_dx = 3 * np.cos(3 * X) * np.cos(5 * Y) - 3.5 * np.sin(7 * X + 2 * Y)
    ex_dy = -5 * np.sin(3 * X) * np.sin(5 * Y) - 1.0 * np.sin(7 * X + 2 * Y)
    for k in range(3):
        assert np.abs(ff[k, :, :g.nx] - f).max() < 1e-13
        assert np.abs(dx[k, :, :g.nx] - ex_dx).max() < 1e-12
        assert np.abs(dy[k, :, :g.nx] - ex_dy).max() < 1e-12
    a = o32.ddx(F)
    b = o32.ddy(F)
    c, d = o32.ddxy(F)
    assert np.abs(a[..., :g.nx] - ex_dx).max() < 1e-12 and np.abs(b[..., :g.nx] - ex_dy).max() < 1e-12
    assert np.array_equal(a, c) and np.array_equal(b, d)


def test_ddz_closed_form():
    cfg = abi.make_config(32, 32, 9)
    o = Oracle(cfg)
    g = cfg.grid
    z = (np.arange(g.nz) + 0.5) * g.dz        # uvp nodes
    f = np.broadcast_to((3 * z + 2 * z * z)[:, None, None], (g.nz, g.ny, g.ld)).copy()
    sentinel = np.full_like(f, 7.0)
    d = o.ddz_uv(f, dfdz=sentinel)
    zw = np.arange(g.nz) * g.dz              # w node jz sits at (jz-1) dz
    exact = 3 + 4 * zw                       # centred difference of a quadratic is exact
    for k in range(1, g.nz - 1):
        assert np.allclose(d[k, :, :g.nx], exact[k], rtol=1e-13)
    assert np.all(d[0] == 7.0)               # level 1 untouched (set by wallstress)
    assert np.all(d[-1] == 0.0)
    fw = np.broadcast_to((3 * zw + 2 * zw * zw)[:, None, None], (g.nz, g.ny, g.ld)).copy()
    d = o.ddz_w(fw)
    for k in range(g.nz - 1):
        assert np.allclose(d[k, :, :g.nx], 3 + 4 * z[k], rtol=1e-13)
    assert np.all(d[-1] == 0.0)


# ---------------- 3/2 padding, dealiasing, test filter ----------------
def test_padd_is_spectral_interpolation():
    """padd + big c2r of a band-limited plane = the same function on the 3/2 grid."""
    cfg = abi.make_config(32, 32, 5)
    o = Oracle(cfg)
    g = cfg.grid
    X, Y = _grid_xy(g)
    Xb, Yb = _grid_xy(g, big=True)
    fun = lambda x, y: np.sin(4 * x - 3 * y) + np.cos(11 * y) * np.sin(2 * x) + 0.25
    F = o.rfft2_fwd(_plane(g, fun(X, Y))) / (g.nx * g.ny)
    B = o.padd(F[None])[0]
    fb = o.rfft2_inv(B, big=True)
    assert np.abs(fb[:, :g.nx2] - fun(Xb, Yb)).max() < 1e-13
    # unpadd o padd = identity off the Nyquist row / column, zero on it
    back = o.unpadd(o.padd(F[None]))[0]
    mask = np.ones_like(F, bool)
    mask[:, -1] = False
    mask[g.ny // 2, :] = False
    assert np.array_equal(back[mask], F[mask]) and np.all(back[~mask] == 0)


def test_dealias_product_exact():
    """dealias1 -> product on the 3/2 grid -> dealias2 reproduces the product of two
    band-limited fields exactly when its wavenumbers stay below nx/2 (3/2 rule)."""
    cfg = abi.make_config(32, 32, 4)
    o = Oracle(cfg)
    g = cfg.grid
    X, Y = _grid_xy(g)
    a = np.sin(5 * X + 2 * Y) + np.cos(3 * Y)
    b = np.cos(6 * X - 4 * Y) + 0.5
    ab = _plane(g, np.stack([a, b]))
    big = o.dealias1(ab)
    prod = big[0:1] * big[1:2]
    out = o.dealias2(prod)
    assert np.abs(out[0, :, :g.nx] - a * b).max() < 1e-13
    # a product whose wavenumbers exceed nx/2 loses exactly those modes (no aliasing back)
    c = np.sin(11 * X)
    big = o.dealias1(_plane(g, np.stack([c, c])))
    out = o.dealias2(big[0:1] * big[1:2])
    assert np.abs(out[0, :, :g.nx] - 0.5).max() < 1e-13     # sin^2 = (1 - cos 22x)/2


def test_test_filter_cutoff_modes():
    cfg = abi.make_config(32, 32, 4)
    o = Oracle(cfg)
    g = cfg.grid
    X, Y = _grid_xy(g)
    G2 = o.test_filter_init(2.0)                 # keeps |k| < 8
    low, high = np.cos(3 * X + 4 * Y), np.sin(9 * X)
    out = o.test_filter(_plane(g, np.stack([low + high + 1.0])), G2)
    assert np.abs(out[0, :, :g.nx] - (low + 1.0)).max() < 1e-13
    G4 = o.test_filter_init(4.0)                 # keeps |k| < 4
    out = o.test_filter(_plane(g, np.stack([low + np.cos(2 * Y)])), G4)
    assert np.abs(out[0, :, :g.nx] - np.cos(2 * Y)).max() < 1e-13


# ---------------- divstress ----------------
def test_divstress_closed_form():
    cfg = abi.make_config(32, 32, 9)
    o = Oracle(cfg)
    g = cfg.grid
    X, Y = _grid_xy(g)
    zu = (np.arange(g.nz) + 0.5) * g.dz
    zw = np.arange(g.nz) * g.dz
    tx = np.stack([np.sin(2 * X) * (1 + z) for z in zu])
    ty = np.stack([np.cos(3 * Y) * (2 - z) for z in zu])
    tzw = np.stack([np.full_like(X, z * z) for z in zw])         # on w-nodes
    d = o.divstress_uv(_plane(g, tx), _plane(g, ty), _plane(g, tzw))
    for k in range(g.nz - 1):
        ex = 2 * np.cos(2 * X) * (1 + zu[k]) - 3 * np.sin(3 * Y) * (2 - zu[k]) + (zw[k + 1] ** 2 - zw[k] ** 2) / g.dz
        assert np.abs(d[k, :, :g.nx] - ex).max() < 1e-11
        assert np.all(d[k, :, g.nx:] == 0)
    assert np.all(d[-1] == abi.BOGUS)
    tzu = np.stack([np.full_like(X, z * z) for z in zu])
    d = o.divstress_w(_plane(g, tx), _plane(g, ty), _plane(g, tzu))
    ex0 = 2 * np.cos(2 * X) * (1 + zu[0]) - 3 * np.sin(3 * Y) * (2 - zu[0])    # no d/dz at the wall
    assert np.abs(d[0, :, :g.nx] - ex0).max() < 1e-11
    for k in range(1, g.nz - 1):
        ex = 2 * np.cos(2 * X) * (1 + zu[k]) - 3 * np.sin(3 * Y) * (2 - zu[k]) + (zu[k] ** 2 - zu[k - 1] ** 2) / g.dz
        assert np.abs(d[k, :, :g.nx] - ex).max() < 1e-11


# ---------------- tridiagonal solve ----------------
def test_tridag_dense_and_tri_debug():
    cfg = abi.make_config(32, 16, 9)
    o = Oracle(cfg)
    g = cfg.grid
    n = g.nz + 1
    rng = np.random.default_rng(1)
    shp = (n, g.ny, g.lh)
    a = rng.uniform(-1, 1, shp)
    c = rng.uniform(-1, 1, shp)
    b = 3.0 + rng.uniform(0, 1, shp)
    r = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    u = o.tridag_array(a, b, c, r)
    for jy in range(g.ny):
        for jx in range(g.lh - 1):
            if jy == g.ny // 2 or (jx == 0 and jy == 0):
                continue
            M = np.diag(b[:, jy, jx]) + np.diag(a[1:, jy, jx], -1) + np.diag(c[:-1, jy, jx], 1)
            assert np.allclose(u[:, jy, jx], np.linalg.solve(M, r[:, jy, jx]), rtol=1e-12, atol=1e-12)
    # TRI_DEBUG known answer (press_stag_array.f90:141-188,245-254): a=1, b=2, c=1,
    # r_j = j; solution of the Toeplitz system tridiag(1,2,1) x = (1..n)
    a = np.ones(shp); b = 2 * np.ones(shp); c = np.ones(shp)
    r = np.broadcast_to(np.arange(1, n + 1, dtype=float)[:, None, None], shp).astype(complex)
    u = o.tridag_array(a, b, c, r)
    # closed form: x_j = j(n+1-j)... no: solve the recurrence x_{j-1} + 2 x_j + x_{j+1} = j
    M = np.diag(2 * np.ones(n)) + np.diag(np.ones(n - 1), -1) + np.diag(np.ones(n - 1), 1)
    exact = np.linalg.solve(M, np.arange(1, n + 1, dtype=float))
    assert np.allclose(u[:, 1, 1].real, exact, rtol=1e-12)
    with pytest.raises(RuntimeError, match="pivot"):
        o.tridag_array(a, np.zeros(shp), c, r)


# ---------------- pressure projection: manufactured solution ----------------
def test_press_stag_array_manufactured():
    """u* = u_df + 1.5 dt grad_d p with a discretely divergence-free u_df: the Poisson
    solve must return p (gauge: plane mean of level 0 is 0) and project() must return
    u_df (press_stag_array.f90:1-361, tridag_array.f90, forcing.f90:122-294)."""
    cfg = abi.make_config(32, 32, 17, lbc_mom=abi.LP_LBC_WALL)
    o = Oracle(cfg)
    g = cfg.grid
    X, Y = _grid_xy(g)
    dt = cfg.dt
    nzl = g.nz
    zc = (np.arange(nzl + 1) - 0.5) * g.dz      # p level j at uvp height (j-1/2) dz, j = 0..nz
    def gz(z):
        return np.cos(np.pi * z / g.L_z) + 0.3 * np.cos(2 * np.pi * z / g.L_z)
    p = np.zeros((nzl + 1, g.ny, g.nx))
    for j in range(nzl + 1):
        p[j] = np.cos(3 * X + 2 * Y) * gz(zc[j]) + 0.2 * np.sin(5 * Y) * gz(zc[j]) + 0.1 * zc[j] ** 2
    p[0] = p[1]; p[nzl] = p[nzl - 1]                # zero normal gradient at wall and lid
    u_df, v_df, w_df = ic.turbulent_ic(g, log_law=False)
    kx = 3.0
    # grad_d p: spectral x-y derivatives at uvp levels 1..nz-1 (p levels 1..nz-1), w-nodes
    dpx = np.zeros((nzl, g.ny, g.nx)); dpy = np.zeros_like(dpx); dpz = np.zeros_like(dpx)
    for k in range(nzl - 1):
        j = k + 1
        dpx[k] = -3 * np.sin(3 * X + 2 * Y) * gz(zc[j])
        dpy[k] = -2 * np.sin(3 * X + 2 * Y) * gz(zc[j]) + np.cos(5 * Y) * gz(zc[j])
    for k in range(nzl):                              # w-node jz = k+1 between p(k) and p(k+1)
        dpz[k] = (p[k + 1] - p[k]) / g.dz
    us = u_df.copy(); vs = v_df.copy(); ws = w_df.copy()
    us[:-1, :, :g.nx] += 1.5 * dt * dpx[:-1]
    vs[:-1, :, :g.nx] += 1.5 * dt * dpy[:-1]
    ws[:-1, :, :g.nx] += 1.5 * dt * dpz[:-1]
    o.set_field("u", us); o.set_field("v", vs); o.set_field("w", ws)
    o.set_field("divtz", np.zeros_like(us))
    o.call("press_stag_array")
    got = o.field("p")[:, :, :g.nx]
    gauge = got[0].mean() - p[0].mean()
    for j in range(nzl):
        assert np.abs(got[j] - gauge - p[j]).max() < 1e-9 * np.abs(p).max(), j
    dpz_num = np.zeros_like(us)
    pp = o.field("p")
    dpz_num[:-1, :, :g.nx] = (pp[1:nzl, :, :g.nx] - pp[:nzl - 1, :, :g.nx]) / g.dz
    o.set_field("dpdz", dpz_num)
    o.call("project")
    for name, ref in (("u", u_df), ("v", v_df), ("w", w_df)):
        a = o.field(name)[:-1, :, :g.nx]
        assert np.abs(a - ref[:-1, :, :g.nx]).max() < 1e-9, name


# ---------------- wall stress, SGS ----------------
def test_wallstress_uniform_closed_form():
    cfg = abi.make_config(32, 32, 9, z0=1e-3)
    o = Oracle(cfg)
    g = cfg.grid
    U, V = 3.0, -1.0
    o.set_field("u", np.full(abi.field_shape(g, "u"), U))
    o.set_field("v", np.full(abi.field_shape(g, "v"), V))
    o.call("wallstress")
    spd = math.hypot(U, V)
    us = 0.4 * spd / math.log(0.5 * g.dz / 1e-3)
    assert np.allclose(o.field("txz")[0, :, :g.nx], -us * us * U / spd, rtol=1e-13)
    assert np.allclose(o.field("tyz")[0, :, :g.nx], -us * us * V / spd, rtol=1e-13)
    assert np.allclose(o.field("dudz")[0, :, :g.nx], us / (0.5 * g.dz * 0.4) * U / spd, rtol=1e-13)


def test_sgs_linear_shear_closed_form():
    """u = S z, Cs^2 = 0.03 (jt = 1, inilag), l = (dx dy dz)^(1/3): txz = -Cs^2 l^2 S^2."""
    cfg = abi.make_config(32, 32, 9)
    o = Oracle(cfg)
    g = cfg.grid
    S = 2.5
    shape = abi.field_shape(g, "u")
    dudz = np.full(shape, S)
    o.set_field("dudz", dudz)
    o.set_time(1)
    o.call("sgs_stag")
    l2 = (g.dx * g.dy * g.dz) ** (2.0 / 3.0)
    for k in range(1, g.nz - 1):
        assert np.allclose(o.field("txz")[k, :, :g.nx], -0.03 * l2 * S * S, rtol=1e-13)
        assert np.allclose(o.field("Nu_t")[k, :, :g.nx], 0.03 * l2 * S, rtol=1e-13)
        assert np.allclose(o.field("txx")[k, :, :g.nx], 0.0)
    assert np.all(o.field("txz")[-1] == 0.0)


def _lag_state(cfg, shift=0.0, scale=1.0):
    o = Oracle(cfg)
    g = cfg.grid
    u, v, w = ic.turbulent_ic(g, log_law=False)
    u = scale * u; v = scale * v; w = scale * w
    u[..., :g.nx] += shift
    o.set_field("u", u); o.set_field("v", v); o.set_field("w", w)
    for comp in "uvw":
        f, dx, dy = o.filt_da(o.field(comp))
        o.set_field(comp, f); o.set_field(f"d{comp}dx", dx); o.set_field(f"d{comp}dy", dy)
    o.set_field("dudz", o.ddz_uv(o.field("u"), dfdz=o.field("dudz")))
    o.set_field("dvdz", o.ddz_uv(o.field("v"), dfdz=o.field("dvdz")))
    o.set_field("dwdz", o.ddz_w(o.field("w")))
    o.set_time(2)
    o.call("sgs_stag")
    return o


def test_lagrangian_galilean_and_scale_invariance():
    """lagrange_Sdep at its initialisation step: L_ij = bar(u_i u_j) - bar(u_i) bar(u_j)
    and M_ij, N_ij depend on velocity differences only, so adding a uniform velocity
    leaves every Lagrangian average and Cs^2 unchanged.  Scaling u by 2 scales LM and MM
    by 16; at initialisation F_LM = eps LM with eps fixed by the 1e-24 floor, F_MM = MM,
    so both scale by 16 and Cs^2_2delta = F_LM/F_MM is scale free (Cs^2 itself is not:
    the 4-delta time scale depends on NN).  A dropped product term, a wrong filter or a
    wrong power breaks one of these."""
    cfg = abi.make_config(32, 32, 17, fp32_casts=0)
    g = cfg.grid
    base = _lag_state(cfg)
    sh = _lag_state(cfg, shift=3.0)
    sc = _lag_state(cfg, scale=2.0)
    lev = slice(1, g.nz - 1)
    f = lambda o, n: o.field(n)[lev, :, :g.nx]
    for n in ("F_LM", "F_MM", "F_QN", "F_NN", "Cs_opt2"):
        a, b = f(base, n), f(sh, n)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), n
    for n in ("F_LM", "F_MM"):
        a, b = f(base, n), f(sc, n)
        pos = a > 1e-20
        assert pos.mean() > 0.05
        assert np.allclose(b[pos], 16.0 * a[pos], rtol=1e-10), n
    assert np.abs(f(sc, "F_NN") - 16.0 * f(base, "F_NN")).max() <= 1e-10 * np.abs(f(base, "F_NN")).max()
    # the top level has beta = 1 (lagrange_Sdep.f90:425-428)
    assert np.all(base.field("Beta")[-1, :, :g.nx] == 1.0)


def test_interpolag_linear_field_exact():
    """trilinear interpolation reproduces a linear field shifted by -u dt exactly away
    from the periodic seam and the walls (interpolag_Sdep.f90:263-365)."""
    cfg = abi.make_config(32, 32, 17)
    o = Oracle(cfg)
    g = cfg.grid
    ii = np.arange(g.nx)[None, None, :]
    jj = np.arange(g.ny)[None, :, None]
    kk = np.arange(g.nz)[:, None, None]
    F = (1.0 + 0.5 * ii + 0.25 * jj + 2.0 * kk) * np.ones((g.nz, g.ny, g.nx))
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        o.set_field(n, _plane(g, F))
    U, V, W = 300.0, -200.0, 50.0          # |x_p| = U dt/dx < 1
    o.set_field("u_lag", np.full(abi.field_shape(g, "u"), U))
    o.set_field("v_lag", np.full(abi.field_shape(g, "u"), V))
    o.set_field("w_lag", np.full(abi.field_shape(g, "u"), W))
    o.call("interpolag_sdep")
    xp, yp, zp = -U * cfg.dt / g.dx, -V * cfg.dt / g.dy, -W * cfg.dt / g.dz
    ex = F + 0.5 * xp + 0.25 * yp + 2.0 * zp
    got = o.field("F_LM")[:, :, :g.nx]
    sl = (slice(2, g.nz - 2), slice(2, g.ny - 2), slice(2, g.nx - 2))
    assert np.abs(got[sl] - ex[sl]).max() < 1e-12
    assert np.all(o.field("u_lag") == 0.0)
    # zero velocity: identity
    o.set_field("F_LM", _plane(g, F))
    for n in ("u_lag", "v_lag", "w_lag"):
        o.set_field(n, np.zeros(abi.field_shape(g, "u")))
    o.call("interpolag_sdep")
    assert np.array_equal(o.field("F_LM")[:, :, :g.nx], F)


# ---------------- convective term ----------------
def test_convec_closed_form():
    """u = A sin(3y) (x-velocity only): omega_z = -3A cos(3y), so c_y = u omega_z =
    -(3A^2/2) sin(6y) (in band, no aliasing with the 3/2 rule); c_x = c_z = 0."""
    cfg = abi.make_config(32, 32, 9)
    o = Oracle(cfg)
    g = cfg.grid
    X, Y = _grid_xy(g)
    A = 1.7
    u = np.stack([A * np.sin(3 * Y)] * g.nz)
    o.set_field("u", _plane(g, u))
    o.set_field("dudy", _plane(g, np.stack([3 * A * np.cos(3 * Y)] * g.nz)))
    o.call("convec")
    for k in range(g.nz - 1):
        assert np.abs(o.field("RHSy")[k, :, :g.nx] - (-(3 * A * A / 2) * np.sin(6 * Y))).max() < 1e-12
        assert np.abs(o.field("RHSx")[k, :, :g.nx]).max() < 1e-12
        assert np.abs(o.field("RHSz")[k, :, :g.nx]).max() < 1e-12
    assert np.all(o.field("RHSx")[-1] == abi.BOGUS)


def test_convec_vertical_staggering():
    """u = a(z) uniform in x-y, dudz on w-nodes: c_z(jz) = -avg_z(u) dudz, i.e. the
    discrete form of -d(u^2/2)/dz (convec.f90:263-279)."""
    cfg = abi.make_config(32, 32, 9)
    o = Oracle(cfg)
    g = cfg.grid
    zu = (np.arange(g.nz) + 0.5) * g.dz
    a = 1.0 + zu + zu ** 2
    o.set_field("u", np.broadcast_to(a[:, None, None], abi.field_shape(g, "u")).copy())
    dudz = np.zeros(g.nz)
    dudz[1:] = (a[1:] - a[:-1]) / g.dz
    o.set_field("dudz", np.broadcast_to(dudz[:, None, None], abi.field_shape(g, "u")).copy())
    o.call("convec")
    for k in range(1, g.nz - 1):
        ex = -(a[k] ** 2 - a[k - 1] ** 2) / (2 * g.dz)
        assert np.allclose(o.field("RHSz")[k, :, :g.nx], ex, rtol=1e-12)
    assert np.allclose(o.field("RHSz")[0, :, :g.nx], 0.0)


# ---------------- whole step: invariants ----------------
def _kill_nyquist(g, f):
    F = np.fft.rfft2(f[..., :g.nx], axes=(-2, -1))
    F[..., -1] = 0
    F[..., g.ny // 2, :] = 0
    out = np.zeros_like(f)
    out[..., :g.nx] = np.fft.irfft2(F, s=(g.ny, g.nx), axes=(-2, -1))
    return out


def test_step_projected_field_is_divergence_free():
    """After press_stag_array + project the interior uvp levels are discretely
    divergence-free (up to the Nyquist modes, which the pressure solve skips and the
    next filt_da removes).  The first and last levels are not: the wall and lid rows
    impose dp/dz = -<divtz> (press_stag_array.f90:149-190) while project keeps w = 0
    there (forcing.f90:215-292)."""
    cfg = abi.make_config(32, 32, 17)
    o = Oracle(cfg)
    g = cfg.grid
    u, v, w = ic.turbulent_ic(g)
    o.set_field("u", u); o.set_field("v", v); o.set_field("w", w)
    o.step(3)
    uu, vv, ww = (_kill_nyquist(g, o.field(n)) for n in "uvw")
    d = ic.discrete_divergence(g, uu, vv, ww)
    scale = np.abs(o.field("dudx")[:, :, :g.nx]).max()
    assert np.abs(d[1:g.nz - 2]).max() < 1e-11 * scale
    assert np.all(o.field("w")[0] == 0) and np.all(o.field("w")[-1] == 0)


def test_timestep_conditions_closed_form():
    cfg = abi.make_config(32, 32, 9)
    o = Oracle(cfg)
    g = cfg.grid
    o.set_field("u", np.full(abi.field_shape(g, "u"), 3.0))
    o.set_field("v", np.full(abi.field_shape(g, "u"), 4.0))
    o.set_field("Nu_t", np.full(abi.field_shape(g, "u"), 0.01))
    cfl, visc = o.timestep_conditions()
    delta = min(g.dx, g.dy, g.dz)
    assert math.isclose(cfl, 5.0 * cfg.dt / delta, rel_tol=1e-14)
    assert math.isclose(visc, cfg.dt * 0.01 / delta ** 2, rel_tol=1e-14)


# ---------------- avg_stats plane statistics: closed form ----------------
def test_plane_stats_closed_form():
    """io.f90:1782-1849 on u = U_k + a_k cos x, v = V_k + b_k sin y, w = W_k + c_k cos x:
    the plane means are U_k, V_k, W_k; on the w-node of level k >= 1, u' = A cos x with
    A = (a_k + a_{k-1})/2 (the average of the two levels), v' = B sin y likewise and
    w' = c_k cos x, so <u'^2> = A^2/2, <v'^2> = B^2/2, <w'^2> = c_k^2/2, <u'w'> = A c_k/2;
    all four vanish on the first level (serial branch).  Amplitudes differ per level, so
    a missing w-node average is caught."""
    from oracle import Oracle
    cfg = abi.make_config(16, 16, 7)
    g = cfg.grid
    o = Oracle(cfg)
    x = np.arange(g.nx) * g.dx
    y = np.arange(g.ny) * g.dy
    a = np.linspace(0.7, 1.3, g.nz)
    b = np.linspace(0.3, 0.1, g.nz)
    c = np.linspace(0.5, 0.9, g.nz)
    U = np.linspace(1.0, 2.0, g.nz)
    V = np.linspace(-0.5, 0.25, g.nz)
    W = np.linspace(0.1, 0.4, g.nz)
    f = {k: np.zeros((g.nz, g.ny, g.ld)) for k in ("u", "v", "w", "Cs_opt2", "Nu_t", "txz")}
    f["u"][:, :, :g.nx] = U[:, None, None] + a[:, None, None] * np.cos(x)[None, None, :]
    f["v"][:, :, :g.nx] = V[:, None, None] + b[:, None, None] * np.sin(y)[None, :, None]
    f["w"][:, :, :g.nx] = W[:, None, None] + c[:, None, None] * np.cos(x)[None, None, :]
    f["Cs_opt2"][:, :, :g.nx] = 0.01 * np.arange(g.nz)[:, None, None]
    f["Nu_t"][:, :, :g.nx] = 2.0
    f["txz"][:, :, :g.nx] = -np.arange(g.nz)[:, None, None]
    f["u"][:, :, g.nx:] = 99.0  # padding columns stay out of the means
    for k, arr in f.items():
        o.set_field(k, arr)
    s = o.plane_stats()
    nl = g.nz - 1
    tol = dict(rtol=1e-13, atol=1e-13)
    assert np.allclose(s[:, 0], U[:nl], **tol) and np.allclose(s[:, 1], V[:nl], **tol)
    assert np.allclose(s[:, 2], W[:nl], **tol)
    assert np.allclose(s[:, 3], 0.01 * np.arange(nl), **tol) and np.allclose(s[:, 4], 2.0, **tol)
    assert np.allclose(s[:, 9], -np.arange(nl), **tol)
    assert np.all(s[0, 5:9] == 0.0)
    A = 0.5 * (a[1:nl] + a[:nl - 1])
    B = 0.5 * (b[1:nl] + b[:nl - 1])
    C_ = c[1:nl]
    assert np.allclose(s[1:, 5], A * A / 2, **tol) and np.allclose(s[1:, 6], B * B / 2, **tol)
    assert np.allclose(s[1:, 7], C_ * C_ / 2, **tol) and np.allclose(s[1:, 8], A * C_ / 2, **tol)
    o.close()


def test_spectrum_closed_form():
    """io.f90:1643-1679 on u = U_k + a_k(y) cos(m1 x + phi_y) + b_k sin(m2 x) + d_k (-1)^i:
    the transform of a row over nx is U_k at m = 0, a e^{i phi}/2 at m1 and b/(2i) at m2, so
    E(0) = U_k^2/2, E(m1) = <a_k^2>_y / 4, E(m2) = b_k^2/4 and every other m < nx/2 is zero:
    the Nyquist wave d_k (-1)^i is left out, and the padding columns are not read.  A lost
    1/nx, a missing 1/2 on the mean, a missing y-average or a Nyquist term each fail."""
    cfg = abi.make_config(16, 8, 6)
    g = cfg.grid
    o = Oracle(cfg)
    i = np.arange(g.nx)
    x = 2 * np.pi * i / g.nx
    m1, m2 = 3, 5
    U = np.linspace(0.5, 2.0, g.nz)
    b = np.linspace(0.2, 0.9, g.nz)
    d = np.linspace(1.0, 3.0, g.nz)
    a = 0.3 + 0.1 * np.arange(g.nz)[:, None] + 0.05 * np.arange(g.ny)[None, :]  # (nz, ny)
    phi = 0.4 * np.arange(g.ny)
    u = np.full((g.nz, g.ny, g.ld), 77.0)
    u[:, :, :g.nx] = (U[:, None, None] + a[:, :, None] * np.cos(m1 * x[None, None, :] + phi[None, :, None])
                      + b[:, None, None] * np.sin(m2 * x)[None, None, :]
                      + d[:, None, None] * ((-1.0) ** i)[None, None, :])
    o.set_field("u", u)
    E = o.spectrum("u")
    nl = g.nz - 1
    want = np.zeros((nl, g.nx // 2))
    want[:, 0] = 0.5 * U[:nl] ** 2
    want[:, m1] = np.mean(a[:nl] ** 2, axis=1) / 4
    want[:, m2] = b[:nl] ** 2 / 4
    assert E.shape == (nl, g.nx // 2)
    assert np.allclose(E, want, rtol=1e-13, atol=1e-15)
    o.close()


def test_spectrum_parseval_and_bins():
    """Parseval on random rows: 2 sum_m E(m) + <X_{nx/2}^2>_y = <u^2> per level, with
    X_{nx/2} = mean_x u (-1)^i the left-out Nyquist coefficient; the bin argument picks that
    bin's volume of PCon, and a field that is no volume (p) or a bin past npcon is refused."""
    from population_balance_amd import droplets
    cfg = abi.make_config(24, 6, 5)
    ds = droplets.DropletSet(3)
    ds.fill(cfg)
    g = cfg.grid
    o = Oracle(cfg)
    rng = np.random.default_rng(17)
    pc = rng.standard_normal((3, g.nz, g.ny, g.ld))
    o.set_field("PCon", pc)
    sgn = (-1.0) ** np.arange(g.nx)
    nl = g.nz - 1
    for bn in range(3):
        E = o.spectrum("PCon", bin=bn)
        r = pc[bn, :nl, :, :g.nx]
        nyq = np.mean(np.mean(r * sgn, axis=2) ** 2, axis=1)
        assert np.allclose(2 * E.sum(axis=1) + nyq, np.mean(r ** 2, axis=(1, 2)), rtol=1e-13)
    with pytest.raises(ValueError):
        o.spectrum("PCon", bin=3)
    with pytest.raises(ValueError):
        o.spectrum("p")
    o.close()
