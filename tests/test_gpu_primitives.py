"""GPU parity of the spectral / z primitives against the binary64 oracle, through the
C ABI (host arrays staged by the library).  Tolerance 1e-12 relative L2 in FP64."""
import numpy as np
import pytest

from conftest import rel_l2
from harness import smooth_random
from population_balance_amd import abi

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", params=[(32, 32, 9), (64, 64, 6), (128, 128, 5), (256, 256, 4)],
                ids=lambda p: f"{p[0]}x{p[1]}x{p[2]}")
def pair(request):
    from oracle import Oracle
    from population_balance_amd.lesgo import Lesgo
    nx, ny, nz = request.param
    cfg = abi.make_config(nx, ny, nz)
    o = Oracle(cfg)
    lg = Lesgo(cfg)
    print(f"\n[case] primitives {nx}x{ny}x{nz}", flush=True)
    yield cfg, o, lg
    lg.close()
    o.close()


def test_wavenumbers_and_filters(pair):
    cfg, o, lg = pair
    for a, b in zip(lg.init_wavenumber(), o.init_wavenumber()):
        assert np.array_equal(a, b)
    for alpha in (2.0, 4.0):
        assert np.array_equal(lg.test_filter_init(alpha), o.test_filter_init(alpha))


def test_filt_da(pair):
    cfg, o, lg = pair
    rng = np.random.default_rng(5)
    g = cfg.grid
    f = rng.standard_normal((g.nz, g.ny, g.ld))
    ref = o.filt_da(f)
    got = [f.copy(), np.zeros_like(f), np.zeros_like(f)]
    lg.filt_da(*got)
    for a, b in zip(got, ref):
        assert rel_l2(a, b, g.nx) < TOL


def test_ddx_ddy_ddxy(pair):
    cfg, o, lg = pair
    g = cfg.grid
    f = smooth_random(cfg, seed=7) + np.random.default_rng(8).standard_normal((g.nz, g.ny, g.ld))
    dx, dy = np.zeros_like(f), np.zeros_like(f)
    lg.ddx(dx, f)
    lg.ddy(dy, f)
    assert rel_l2(dx, o.ddx(f), g.nx) < TOL
    assert rel_l2(dy, o.ddy(f), g.nx) < TOL
    a, b = np.zeros_like(f), np.zeros_like(f)
    lg.ddxy(a, b, f)
    ra, rb = o.ddxy(f)
    assert rel_l2(a, ra, g.nx) < TOL and rel_l2(b, rb, g.nx) < TOL


def test_ddz(pair):
    cfg, o, lg = pair
    g = cfg.grid
    rng = np.random.default_rng(9)
    f = rng.standard_normal((g.nz, g.ny, g.ld))
    d0 = rng.standard_normal((g.nz, g.ny, g.ld))
    d = d0.copy()
    lg.ddz_uv(d, f)
    ref = o.ddz_uv(f, dfdz=d0)
    assert np.array_equal(d[..., :g.nx], ref[..., :g.nx])
    d = np.zeros_like(f)
    lg.ddz_w(d, f)
    assert np.array_equal(d[..., :g.nx], o.ddz_w(f)[..., :g.nx])


def test_padd_unpadd(pair):
    cfg, o, lg = pair
    g = cfg.grid
    rng = np.random.default_rng(10)
    u = rng.standard_normal((3, g.ny, g.lh)) + 1j * rng.standard_normal((3, g.ny, g.lh))
    big = np.zeros((3, g.ny2, g.lh_big), complex)
    lg.padd(big, u)
    assert np.array_equal(big, o.padd(u))
    cc = np.zeros_like(u)
    lg.unpadd(cc, big)
    assert np.array_equal(cc, o.unpadd(big))


def test_dealias(pair):
    cfg, o, lg = pair
    g = cfg.grid
    rng = np.random.default_rng(11)
    u = rng.standard_normal((g.nz, g.ny, g.ld))
    big = np.zeros((g.nz, g.ny2, g.ld_big))
    lg.dealias1(u, big)
    ref = o.dealias1(u)
    assert rel_l2(big, ref, g.nx2) < TOL
    out = np.zeros_like(u)
    lg.dealias2(out, ref.copy())
    assert rel_l2(out, o.dealias2(ref), g.nx) < TOL


def test_test_filter(pair):
    cfg, o, lg = pair
    g = cfg.grid
    rng = np.random.default_rng(12)
    f = rng.standard_normal((g.nz, g.ny, g.ld))
    for alpha in (2.0, 4.0):
        G = o.test_filter_init(alpha)
        got = f.copy()
        lg.test_filter(got, G)
        assert rel_l2(got, o.test_filter(f, G), g.nx) < TOL


def test_divstress(pair):
    cfg, o, lg = pair
    g = cfg.grid
    rng = np.random.default_rng(13)
    tx, ty, tz = (rng.standard_normal((g.nz, g.ny, g.ld)) for _ in range(3))
    d = np.zeros_like(tx)
    lg.divstress_uv(d, tx, ty, tz)
    ref = o.divstress_uv(tx, ty, tz)
    assert rel_l2(d[:-1], ref[:-1], g.nx) < TOL
    assert np.all(d[-1] == abi.BOGUS)
    lg.divstress_w(d, tx, ty, tz)
    assert rel_l2(d, o.divstress_w(tx, ty, tz), g.nx) < TOL


def test_tridag_array(pair):
    cfg, o, lg = pair
    g = cfg.grid
    n = g.nz + 1
    rng = np.random.default_rng(14)
    shp = (n, g.ny, g.lh)
    a = rng.uniform(0.5, 1.0, shp)
    c = rng.uniform(0.5, 1.0, shp)
    b = -(a + c) - rng.uniform(0.1, 1.0, shp)
    r = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    u = np.zeros(shp, complex)
    lg.tridag_array(a, b, c, r, u)
    ref = o.tridag_array(a, b, c, r)
    m = np.ones((g.ny, g.lh), bool)
    m[g.ny // 2, :] = False
    m[:, -1] = False
    m[0, 0] = False
    assert rel_l2(u[:, m], ref[:, m]) < TOL


def test_tridag_array_zero_pivot(pair):
    """A zero leading pivot is reported as the reference's stop (tridag_array.f90:35-38):
    LP_ERR_PIVOT, raised from the host; the flag is cleared, so a regular system solved
    next reports no error."""
    from population_balance_amd.lesgo import LesgoError
    cfg, o, lg = pair
    g = cfg.grid
    n = g.nz + 1
    shp = (n, g.ny, g.lh)
    a = np.ones(shp)
    c = np.ones(shp)
    r = np.ones(shp, complex)
    u = np.zeros(shp, complex)
    with pytest.raises(LesgoError, match="pivot"):
        lg.tridag_array(a, np.zeros(shp), c, r, u)
    lg.tridag_array(a, -3.0 * np.ones(shp), c, r, u)
    assert np.all(np.isfinite(u))
