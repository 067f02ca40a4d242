"""The oracle's discrete adjoint of the frozen-coefficient step (CPU), pinned by central
finite differences of the frozen-coefficient forward step: for random adjoint seeds
(a, r) and a random direction d of the state (u, v, w and RHS_f),
    <adjoint(a, r), d>  ==  d/de [ <a, u(x + e d)> + <r, RHS(x + e d)> ]
to the truncation error of the central difference.  The frozen step is quadratic in the
state, so the central difference is exact up to rounding."""
import ctypes as C

import numpy as np
import pytest

from oracle import Oracle
from population_balance_amd import abi, ic

VEL = ("u", "v", "w")
RHS = ("RHSx", "RHSy", "RHSz")
RHSF = ("RHSx_f", "RHSy_f", "RHSz_f")


def _bind(o):
    L = o.L
    dp = C.POINTER(C.c_double)
    L.or_set_frozen.argtypes = [C.c_void_p, dp, dp]
    L.or_set_frozen.restype = None
    L.or_step_adjoint.argtypes = [C.c_void_p, dp, dp, dp, dp]
    L.or_step_adjoint.restype = C.c_int
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _phys(g, a):
    out = np.zeros_like(a)
    out[..., :g.nx] = a[..., :g.nx]
    return out


def _base(cfg, steps_before, seed=5):
    """a base state after `steps_before` ordinary steps, with the Nu_t and wall u* of
    the next (ordinary) step recorded for freezing"""
    o = Oracle(cfg)
    g = cfg.grid
    u, v, w = ic.turbulent_ic(g, seed=seed)
    for n, a in zip(VEL, (u, v, w)):
        o.set_field(n, a)
    o.step(steps_before)
    state = {n: o.field(n).copy() for n in VEL + RHS}
    jt = o.time()
    o.step(1)
    nut = o.field("Nu_t").copy()
    ust = o.field("ustar").copy()
    o.close()
    return state, jt, nut, ust


def _frozen_step(cfg, state, jt, nut, ust):
    o = Oracle(cfg)
    L = _bind(o)
    for n in VEL + RHS:
        o.set_field(n, state[n])
    o.set_time(*jt)
    L.or_set_frozen(o.h, _p(nut), _p(ust))
    o.step(1)
    out = {n: o.field(n).copy() for n in VEL + RHS}
    L.or_set_frozen(o.h, None, None)
    o.close()
    return out


@pytest.mark.parametrize("steps_before,kw", [
    (0, {}),                                            # jt = 1: Euler start, wall law
    (2, {}),                                            # AB2 with RHS_f
    (2, {"use_mean_p_force": 1, "mean_p_force": 0.7}),  # top level copied from below
    (2, {"lbc_mom": abi.LP_LBC_STRESS_FREE}),           # stress-free bottom
])
def test_step_adjoint_against_central_differences(steps_before, kw):
    cfg = abi.make_config(32, 32, 9, **kw)
    g = cfg.grid
    state, jt, nut, ust = _base(cfg, steps_before)
    rng = np.random.default_rng(11 + steps_before)
    shp = abi.field_shape(g, "u")
    seed_u = [_phys(g, rng.standard_normal(shp)) for _ in VEL]
    seed_r = [_phys(g, rng.standard_normal(shp)) for _ in RHS]
    for s in seed_r:
        s[-1] = 0.0  # the RHS the step leaves has nz-1 levels
    d_u = [_phys(g, rng.standard_normal(shp)) for _ in VEL]
    d_r = [_phys(g, rng.standard_normal(shp)) for _ in RHS]

    def J(eps):
        st = dict(state)
        for n, d in zip(VEL, d_u):
            st[n] = state[n] + eps * d
        # the step reads RHS_f = the RHS it is given (main.f90:252-254)
        for n, d in zip(RHS, d_r):
            st[n] = state[n] + eps * d
        out = _frozen_step(cfg, st, jt, nut, ust)
        tot = 0.0
        for n, a in zip(VEL, seed_u):
            tot += np.sum(a[..., :g.nx] * out[n][..., :g.nx])
        for n, a in zip(RHS, seed_r):
            tot += np.sum(a[:g.nz - 1, :, :g.nx] * out[n][:g.nz - 1, :, :g.nx])
        return tot

    # adjoint at the base state
    o = Oracle(cfg)
    L = _bind(o)
    for n in VEL + RHS:
        o.set_field(n, state[n])
    o.set_time(*jt)
    ub = np.ascontiguousarray(np.stack(seed_u))
    rb = np.ascontiguousarray(np.stack(seed_r))
    assert L.or_step_adjoint(o.h, _p(nut), _p(ust), _p(ub), _p(rb)) == 0
    o.close()
    adj = sum(np.sum(ub[q][..., :g.nx] * d_u[q][..., :g.nx]) for q in range(3))
    adj += sum(np.sum(rb[q][..., :g.nx] * d_r[q][..., :g.nx]) for q in range(3))
    # the frozen step is quadratic in the state, so a central difference with a large
    # step is exact up to rounding, and small paths (the wall law, the mean mode, the lid
    # row) show above the tolerance
    eps = 0.5
    fd = (J(eps) - J(-eps)) / (2 * eps)
    print(f"adjoint {adj:.15e}  central difference {fd:.15e}  rel {abs(adj - fd) / abs(fd):.2e}")
    assert abs(adj - fd) <= 1e-12 * abs(fd)
    if steps_before == 0:
        assert np.all(rb == 0.0)  # the Euler start does not read RHS_f
