"""The oracle's discrete adjoint of the frozen-coefficient step with droplets (CPU): the
velocity and droplet adjoints together, pinned by central differences of the
frozen-coefficient forward step.  For random adjoint seeds of the step's outputs
(u, v, w, RHS, PCon, RHS_PCon) and a random direction of its inputs,
    <adjoint(seeds), direction> == d/de <seeds, step(state + e direction)>.
The droplet transport is linear in C on each branch of the SMART limiter and bilinear in
(face velocity, face value), so with a small step the central difference is exact up to
rounding, provided no face changes limiter branch inside +-e: the directions of C are
relative (C times a random factor), since an absolute one would flip the branches in the
source's far tails, where neighbouring cells differ by less than the perturbation."""
import ctypes as C

import numpy as np
import pytest

from oracle import Oracle
from population_balance_amd import abi, droplets, ic

VEL = ("u", "v", "w")
RHS = ("RHSx", "RHSy", "RHSz")


def _bind(o):
    L = o.L
    dp = C.POINTER(C.c_double)
    L.or_set_frozen.argtypes = [C.c_void_p, dp, dp]
    L.or_set_frozen.restype = None
    L.or_step_adjoint2.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp]
    L.or_step_adjoint2.restype = C.c_int
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _phys(g, a):
    out = np.zeros_like(a)
    out[..., :g.nx] = a[..., :g.nx]
    return out


def _cfg(bc, active=True, settling=True, npcon=3):
    cfg = abi.make_config(24, 24, 11)
    ds = droplets.DropletSet(npcon)
    ds.fill(cfg, bc=bc, active=active, settling=settling, inertia=False, pbe=True)
    return cfg, ds


def _base(cfg, ds, steps_before, seed=5):
    o = Oracle(cfg)
    g = cfg.grid
    u, v, w = ic.turbulent_ic(g, seed=seed)
    for n, a in zip(VEL, (u, v, w)):
        o.set_field(n, a)
    rng = np.random.default_rng(seed)
    c0 = droplets.initial_concentration(cfg, ds) * (1.0 + 0.3 * rng.random(abi.field_shape(g, "PCon", ds.diameter.size)))
    o.set_field("PCon", _phys(g, c0))
    o.step(steps_before)
    names = VEL + RHS + ("PCon", "RHS_PCon")
    state = {n: o.field(n).copy() for n in names}
    jt = o.time()
    o.step(1)
    nut, ust = o.field("Nu_t").copy(), o.field("ustar").copy()
    o.close()
    return state, jt, nut, ust


def _frozen_step(cfg, state, jt, nut, ust):
    o = Oracle(cfg)
    L = _bind(o)
    for n, a in state.items():
        o.set_field(n, a)
    o.set_time(*jt)
    L.or_set_frozen(o.h, _p(nut), _p(ust))
    o.step(1)
    out = {n: o.field(n).copy() for n in state}
    L.or_set_frozen(o.h, None, None)
    o.close()
    return out


@pytest.mark.parametrize("bc,steps_before,active", [
    ((0, 0), 0, True),    # periodic, the droplet step's Euler start
    ((0, 0), 2, True),    # periodic, AB2
    ((1, 1), 2, True),    # inlet / outlet ghosts
    ((2, 2), 2, False),   # upwind-sign ghosts, no buoyancy
])
def test_step_adjoint_with_droplets_against_central_differences(bc, steps_before, active):
    cfg, ds = _cfg(bc, active=active)
    g = cfg.grid
    npc = cfg.drop.npcon
    state, jt, nut, ust = _base(cfg, ds, steps_before)
    rng = np.random.default_rng(7 + steps_before)
    shp, shc = abi.field_shape(g, "u"), abi.field_shape(g, "PCon", npc)
    cscale = float(np.abs(state["PCon"]).max())
    seeds = {n: _phys(g, rng.standard_normal(shp)) for n in VEL + RHS}
    for n in RHS:
        seeds[n][-1] = 0.0  # the RHS the step leaves has nz-1 levels
    seeds["PCon"] = _phys(g, rng.standard_normal(shc))
    seeds["RHS_PCon"] = _phys(g, rng.standard_normal(shc))
    seeds["RHS_PCon"][:, -1] = 0.0
    d = {n: _phys(g, rng.standard_normal(shp)) for n in VEL + RHS}
    d["PCon"] = _phys(g, state["PCon"] * rng.standard_normal(shc))
    d["PCon"][:, -1] = d["PCon"][:, -2]  # the top level is a copy of the one below
    d["RHS_PCon"] = _phys(g, cscale * rng.standard_normal(shc))

    def J(eps):
        st = {n: state[n] + eps * d[n] for n in state}
        out = _frozen_step(cfg, st, jt, nut, ust)
        tot = 0.0
        for n in VEL + ("PCon",):
            tot += np.sum(seeds[n][..., :g.nx] * out[n][..., :g.nx])
        for n in RHS:
            tot += np.sum(seeds[n][:g.nz - 1, :, :g.nx] * out[n][:g.nz - 1, :, :g.nx])
        tot += np.sum(seeds["RHS_PCon"][:, :g.nz - 1, :, :g.nx] * out["RHS_PCon"][:, :g.nz - 1, :, :g.nx])
        return tot

    o = Oracle(cfg)
    L = _bind(o)
    for n, a in state.items():
        o.set_field(n, a)
    o.set_time(*jt)
    ub = np.ascontiguousarray(np.stack([seeds[n] for n in VEL]))
    rb = np.ascontiguousarray(np.stack([seeds[n] for n in RHS]))
    cb = np.ascontiguousarray(seeds["PCon"].copy())
    rcb = np.ascontiguousarray(seeds["RHS_PCon"].copy())
    assert L.or_step_adjoint2(o.h, _p(nut), _p(ust), _p(ub), _p(rb), _p(cb), _p(rcb)) == 0
    o.close()
    adj = sum(np.sum(ub[q][..., :g.nx] * d[n][..., :g.nx]) for q, n in enumerate(VEL))
    adj += sum(np.sum(rb[q][..., :g.nx] * d[n][..., :g.nx]) for q, n in enumerate(RHS))
    adj += np.sum(cb[..., :g.nx] * d["PCon"][..., :g.nx]) + np.sum(rcb[..., :g.nx] * d["RHS_PCon"][..., :g.nx])
    eps = 1e-6
    fd = (J(eps) - J(-eps)) / (2 * eps)
    print(f"adjoint {adj:.15e}  central difference {fd:.15e}  rel {abs(adj - fd) / abs(fd):.2e}")
    assert abs(adj - fd) <= 1e-7 * abs(fd)
    if steps_before == 0:
        assert np.all(rcb == 0.0)  # the Euler start reads no RHS_PCon


def test_droplet_adjoint_parts_separately():
    """the C -> C and u -> C parts alone (each a sizeable share of the total), so that a
    dropped transpose in either shows above the tolerance"""
    cfg, ds = _cfg((0, 0))
    g = cfg.grid
    npc = cfg.drop.npcon
    state, jt, nut, ust = _base(cfg, ds, 2)
    rng = np.random.default_rng(3)
    shp, shc = abi.field_shape(g, "u"), abi.field_shape(g, "PCon", npc)
    seedC = _phys(g, rng.standard_normal(shc))
    for part in ("C", "u"):
        d = {n: np.zeros(shp) for n in VEL + RHS}
        d["PCon"], d["RHS_PCon"] = np.zeros(shc), np.zeros(shc)
        if part == "C":
            d["PCon"] = _phys(g, state["PCon"] * rng.standard_normal(shc))
            d["PCon"][:, -1] = d["PCon"][:, -2]
        else:
            d["u"] = _phys(g, rng.standard_normal(shp))
            d["w"] = _phys(g, rng.standard_normal(shp))

        def J(eps):
            st = {n: state[n] + eps * d[n] for n in state}
            out = _frozen_step(cfg, st, jt, nut, ust)
            return np.sum(seedC[..., :g.nx] * out["PCon"][..., :g.nx])

        o = Oracle(cfg)
        L = _bind(o)
        for n, a in state.items():
            o.set_field(n, a)
        o.set_time(*jt)
        ub, rb = np.zeros((3,) + shp), np.zeros((3,) + shp)
        cb, rcb = np.ascontiguousarray(seedC.copy()), np.zeros(shc)
        assert L.or_step_adjoint2(o.h, _p(nut), _p(ust), _p(ub), _p(rb), _p(cb), _p(rcb)) == 0
        o.close()
        adj = sum(np.sum(ub[q][..., :g.nx] * d[n][..., :g.nx]) for q, n in enumerate(VEL))
        adj += np.sum(cb[..., :g.nx] * d["PCon"][..., :g.nx])
        eps = 1e-6
        fd = (J(eps) - J(-eps)) / (2 * eps)
        print(f"{part}: adjoint {adj:.15e}  central difference {fd:.15e}  rel {abs(adj - fd) / abs(fd):.2e}")
        assert abs(adj - fd) <= 1e-7 * abs(fd)


def test_buoyancy_path():
    """C -> RHSz -> w (calcbeta_pcon, main.f90:458-460): with centimetre drops the
    buoyancy is large enough for a central difference to see it"""
    cfg = abi.make_config(24, 24, 11)
    ds = droplets.DropletSet(2, d_lo=5e-2, d_hi=1e-1)
    ds.fill(cfg, bc=(0, 0), active=True, settling=False, inertia=False, pbe=False)
    g = cfg.grid
    state, jt, nut, ust = _base(cfg, ds, 2)
    state["PCon"] = state["PCon"] * 1e4  # O(10) concentrations: a buoyancy of O(1e-2)
    rng = np.random.default_rng(4)
    shp, shc = abi.field_shape(g, "u"), abi.field_shape(g, "PCon", 2)
    seedw = _phys(g, rng.standard_normal(shp))
    d = _phys(g, state["PCon"] * rng.standard_normal(shc))
    d[:, -1] = d[:, -2]

    def J(eps):
        st = dict(state)
        st["PCon"] = state["PCon"] + eps * d
        out = _frozen_step(cfg, st, jt, nut, ust)
        return np.sum(seedw[..., :g.nx] * out["w"][..., :g.nx])

    o = Oracle(cfg)
    L = _bind(o)
    for n, a in state.items():
        o.set_field(n, a)
    o.set_time(*jt)
    ub, rb = np.zeros((3,) + shp), np.zeros((3,) + shp)
    ub[2] = seedw
    cb, rcb = np.zeros(shc), np.zeros(shc)
    assert L.or_step_adjoint2(o.h, _p(nut), _p(ust), _p(ub), _p(rb), _p(cb), _p(rcb)) == 0
    o.close()
    adj = np.sum(cb[..., :g.nx] * d[..., :g.nx])
    eps = 1.0  # w is linear in C (through the buoyancy only): any step is exact
    fd = (J(eps) - J(-eps)) / (2 * eps)
    print(f"buoyancy: adjoint {adj:.15e}  central difference {fd:.15e}")
    assert abs(fd) > 1e-6
    assert abs(adj - fd) <= 1e-6 * abs(fd)
