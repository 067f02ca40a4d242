"""The oracle's other dynamic SGS models (sgs_stag.f90:188-198): std_dynamic.f90 (model 2),
scaledep_dynamic.f90 (model 3), lagrange_Ssim.f90 with interpolag_Ssim.f90 (model 4),
checked on the CPU against what the mathematics fixes:

* the plane sums expand: with M(b) = 2 delta^2 (|S|S_ij bar - 4 b |S|bar S_ij bar),
  sum L.M(b) = n (b a1 - b1) and sum M(b).M(b) = n (c1 - b e1 + b^2 d1) in the reference's
  coefficients a1..e1 (scaledep_dynamic.f90:156-175), so the per-point contraction of
  models 2 (b = 1) and 3 must equal (b a1 - b1) / (c1 - b e1 + b^2 d1);
* model 3's b is (to rtnewt's stopping rule, a Newton step below 0.001) a root of the
  quintic that equates the 2-delta coefficient with the 4-delta one formed with b^2:
  (b a1 - b1)/(c1 - b e1 + b^2 d1) = (b^2 a2 - b2)/(c2 - b^2 e2 + b^4 d2);
* Cs^2 is unchanged by a uniform shift of u, v (Galilean) and, for the plane-averaged
  models, by scaling u, v, w;
* interpolag_Ssim moves the averages by whole cells for a velocity of one cell per step,
  and at the wall doubles a downward zp without the clamp of interpolag_Sdep;
* lagrange_Ssim's first (initialising) step follows from L.M and M.M in closed form."""
import numpy as np
import pytest

from oracle import Oracle
from population_balance_amd import abi, ic

NX, NZ = 32, 9


def _cfg(model, **kw):
    cfg = abi.make_config(NX, NX, NZ, model=model, inilag=0, dyn_init=1, **kw)
    return cfg


def _run(model, u, v, w, **kw):
    """one step from (u, v, w): the SGS model runs on the initial state"""
    cfg = _cfg(model, **kw)
    o = Oracle(cfg)
    for n, a in zip(("u", "v", "w"), (u, v, w)):
        o.set_field(n, a)
    o.step(1)
    return o


def _sgs(o):
    """main.f90:277-327 up to sgs_stag on the oracle's state (the velocities filtered, the
    derivatives, the wall law, the SGS model), leaving u, v, w as the model saw them"""
    for comp in "uvw":
        f2, dx, dy = o.filt_da(o.field(comp).copy())
        o.set_field(comp, f2)
        o.set_field(f"d{comp}dx", dx)
        o.set_field(f"d{comp}dy", dy)
    o.set_field("dudz", o.ddz_uv(o.field("u"), dfdz=o.field("dudz")))
    o.set_field("dvdz", o.ddz_uv(o.field("v"), dfdz=o.field("dvdz")))
    o.set_field("dwdz", o.ddz_w(o.field("w")))
    o.call("wallstress")
    o.call("sgs_stag")


def _at(model, u, v, w, **kw):
    cfg = _cfg(model, **kw)
    o = Oracle(cfg)
    for n, a in zip(("u", "v", "w"), (u, v, w)):
        o.set_field(n, a)
    _sgs(o)
    return o


@pytest.fixture(scope="module")
def state():
    g = abi.make_grid(NX, NX, NZ)
    return ic.turbulent_ic(g, seed=9)


@pytest.mark.parametrize("model", [2, 3])
def test_plane_models_expand(state, model):
    o = _at(model, *state)
    g = o.cfg.grid
    cs = o.field("Cs_opt2")
    for k in (0, 1, 3, NZ - 2, NZ - 1):
        a1, b1, c1, d1, e1, a2, b2, c2, d2, e2, beta, cs_k = o.plane_dynamic_level(model, k)
        assert np.all(cs[k] == cs_k)  # one value over the whole plane
        want = max(0.0, (beta * a1 - b1) / (c1 - beta * e1 + beta ** 2 * d1))
        assert cs_k == pytest.approx(want, rel=1e-11, abs=1e-300), (k, cs_k, want)
        if model == 2:
            assert beta == 1.0
        else:
            assert beta != 1.0  # rtnewt converged
            # the quintic of the equal-coefficient condition, refined from the oracle's b
            # (rtnewt stops once a Newton step is below 0.001): b must lie next to its root,
            # where the 2-delta and 4-delta coefficients agree
            P = np.polynomial.Polynomial([b2 * c1 - b1 * c2, a1 * c2 - b2 * e1, b2 * d1 + b1 * e2 - a2 * c1,
                                          a2 * e1 - a1 * e2, -a2 * d1 - b1 * d2, a1 * d2])
            r = beta
            for _ in range(60):
                r = r - P(r) / P.deriv()(r)
            assert abs(r - beta) < 1e-3, (k, beta, r)
            r2 = (r * a1 - b1) / (c1 - r * e1 + r ** 2 * d1)
            r4 = (r ** 2 * a2 - b2) / (c2 - r ** 2 * e2 + r ** 4 * d2)
            assert r2 == pytest.approx(r4, rel=1e-9), (k, r2, r4)
    o.close()


@pytest.mark.parametrize("model", [2, 3, 4])
def test_galilean_and_scale_invariance(state, model):
    u, v, w = state
    g = abi.make_grid(NX, NX, NZ)

    def cs_of(uu, vv, ww):
        o = _run(model, uu, vv, ww)
        c = o.field("Cs_opt2").copy()
        o.close()
        return c[..., :g.nx]
    c0 = cs_of(u, v, w)
    pad = np.zeros_like(u)
    pad[..., :g.nx] = 1.0
    # levels from 2 up: the wall law (level 1 of the strain) reads the absolute velocity
    cg = cs_of(u + 0.4 * pad, v - 0.3 * pad, w)
    assert np.allclose(cg[1:], c0[1:], rtol=1e-9, atol=1e-30)
    if model == 4:
        return  # the Lagrangian time scale, the 1e-32 floor and REAL(4) casts depend on the size
    cl = cs_of(3.0 * u, 3.0 * v, 3.0 * w)
    assert np.allclose(cl, c0, rtol=1e-9, atol=1e-30)


def _interp_case(ssim, U=0.0, V=0.0, W=0.0, nz=NZ):
    """F_LM, F_MM from a smooth field, lag velocities U, V, W (one step's worth), one
    interpolag call"""
    cfg = abi.make_config(NX, NX, nz, model=4 if ssim else 5)
    g = cfg.grid
    o = Oracle(cfg)
    k = np.arange(g.nz)[:, None, None]
    y = np.arange(g.ny)[None, :, None]
    x = np.arange(g.ld)[None, None, :]
    F = 1.0 + 0.3 * np.sin(2 * np.pi * x / g.nx) * np.cos(2 * np.pi * y / g.ny) + 0.1 * k + 0 * x
    F[..., g.nx:] = 0.0
    o.set_field("F_LM", F)
    o.set_field("F_MM", 2.0 * F)
    for n, a in (("u_lag", U), ("v_lag", V), ("w_lag", W)):
        o.set_field(n, np.full((g.nz, g.ny, g.ld), a))
    o.call("interpolag_ssim" if ssim else "interpolag_sdep")
    out = o.field("F_LM").copy(), o.field("F_MM").copy()
    o.close()
    return F, out, g


def test_interpolag_ssim_whole_cell_shifts():
    """u dt/dx = 1: every average comes from the upstream cell (x - 1, periodic)"""
    g = abi.make_grid(NX, NX, NZ)
    dt = 2e-4
    F, (flm, fmm), g = _interp_case(True, U=g.dx / dt)
    want = np.roll(F[..., :g.nx], 1, axis=2)
    assert np.allclose(flm[..., :g.nx], want, rtol=1e-13, atol=0)
    assert np.allclose(fmm[..., :g.nx], 2 * want, rtol=1e-13, atol=0)
    F, (flm, _), g = _interp_case(True, V=-g.dy / dt)
    assert np.allclose(flm[..., :g.nx], np.roll(F[..., :g.nx], -1, axis=1), rtol=1e-13, atol=0)


def test_interpolag_ssim_wall_rule():
    """w_lag = W on every level: zp = -W dt / dz.  At the first w node (level 2) Ssim doubles
    a downward zp (zp = -1.2 reaches past the wall ghost, an extrapolation), Sdep clamps
    it to -1; at level 1 Sdep copies level 2's, Ssim keeps w_lag(1) = w_lag(2)/4 with a
    downward zp unchanged"""
    g = abi.make_grid(NX, NX, NZ)
    dt = 2e-4
    W = 0.6 * g.dz / dt  # zp = -0.6 away from the wall
    F, (ss, _), g = _interp_case(True, W=W)
    F, (sd, _), g = _interp_case(False, W=W)
    # level 2 (index 1): F(1) - z = 1.2 F(1) back: frac_z = 1 - 1.2, comp_z = 1.2 toward k-1
    f0, f1 = F[0, :, :g.nx], F[1, :, :g.nx]
    assert np.allclose(ss[1, :, :g.nx], (1 - 1.2) * f1 + 1.2 * f0, rtol=1e-12, atol=1e-14)
    assert np.allclose(sd[1, :, :g.nx], 0.0 * f1 + 1.0 * f0, rtol=1e-12, atol=1e-14)
    # level 1 (index 0): zp = -0.6 / 4 (w_lag(1) = w_lag(2)/4), the ghost below repeats level 1
    assert np.allclose(ss[0, :, :g.nx], f0, rtol=1e-13, atol=0)
    # interior levels: a plain 0.6 step down
    k = 4
    assert np.allclose(ss[k, :, :g.nx], 0.4 * F[k, :, :g.nx] + 0.6 * F[k - 1, :, :g.nx], rtol=1e-12, atol=0)


@pytest.mark.parametrize("casts", [0, 1])
def test_lagrange_ssim_initial_step(state, casts):
    """inilag, jt = DYN_init: F_MM = M.M, F_LM = 0.025 M.M, then one relaxation with
    Tn = opftime delta (F_LM F_MM)^(-1/8) (REAL(4) where the reference converts) and
    Cs^2 = F_LM / F_MM; the padding columns hold Cs^2 = 1e-32"""
    u, v, w = state
    cfg = abi.make_config(NX, NX, NZ, model=4, inilag=1, dyn_init=2, cs_count=1, fp32_casts=casts)
    g = cfg.grid
    o = Oracle(cfg)
    for n, a in zip(("u", "v", "w"), (u, v, w)):
        o.set_field(n, a)
    o.step(1)  # jt = 1: Cs^2 = 0.03
    assert np.all(o.field("Cs_opt2") == 0.03)
    _sgs(o)  # jt = 2 = DYN_init: the model's first step
    delta = (g.dx * g.dy * g.dz) ** (1 / 3)
    r4 = (lambda a: np.float32(a).astype(np.float64)) if casts else (lambda a: a)
    tiny = float(np.float32(1e-32)) if casts else 1e-32
    for k in (0, 2, NZ - 1):
        lm, mm = o.dyn_contract(k)  # from the strain of the step just taken
        lm, mm = lm[:, :g.nx], mm[:, :g.nx]
        flm0, fmm0 = 0.025 * mm, mm
        Tn = np.maximum(r4(flm0 * fmm0), tiny)
        Tn = 1.5 * delta * Tn ** (-1 / 8)
        dum = 2e-4 / Tn
        eps = dum / (1 + dum)
        flm = np.maximum(tiny, r4(eps * lm + (1 - eps) * flm0))
        fmm = eps * mm + (1 - eps) * fmm0
        cs = np.maximum(tiny, r4(flm / fmm))
        assert np.allclose(o.field("F_LM")[k, :, :g.nx], flm, rtol=1e-12, atol=0)
        assert np.allclose(o.field("F_MM")[k, :, :g.nx], fmm, rtol=1e-12, atol=0)
        assert np.allclose(o.field("Cs_opt2")[k, :, :g.nx], cs, rtol=1e-12, atol=0)
        assert np.all(o.field("Cs_opt2")[k, :, g.ld - 2:] == tiny)
    o.close()
