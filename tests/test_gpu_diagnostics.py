"""GPU parity of the run diagnostics (energy.f90:1-47 lp_energy; io_crossflow_mod.f90:
2452-2799 lp_tavg_reset / lp_tavg_compute / lp_tavg_get) against the oracle on the same
state: a few GPU steps (with droplet bins), then the GPU's own fields are loaded into the
oracle, so that only the diagnostics are compared.  Binary32 fields of a mixed context
convert to binary64 exactly, so both precisions are held to 1e-13 of each quantity's
largest magnitude."""
import numpy as np
import pytest

from population_balance_amd import abi, droplets, ic

pytestmark = pytest.mark.gpu

TOL = 1e-13
TAVG_IN = ("u", "v", "w", "p", "txx", "tyy", "tzz", "txy", "txz", "tyz", "dudx", "dudy", "dudz", "dvdx",
           "dvdy", "dvdz", "dwdx", "dwdy", "dwdz")


def _run(precision, npc, nsteps=3, nx=32, nz=17):
    from population_balance_amd.lesgo import Lesgo
    cfg = abi.make_config(nx, nx, nz, precision=precision)
    ds = None
    if npc:
        ds = droplets.DropletSet(npc)
        ds.fill(cfg, bc=(0, 0), inertia=False)
    lg = Lesgo(cfg)
    for n, a in zip(("u", "v", "w"), ic.turbulent_ic(cfg.grid, seed=11)):
        lg.upload(n, a)
    if npc:
        C = droplets.initial_concentration(cfg, ds)
        C[..., cfg.grid.nx:] = 0.0
        lg.upload("PCon", C)
    lg.step(nsteps)
    return cfg, ds, lg


def _oracle_of(cfg, lg, npc):
    from oracle import Oracle
    o = Oracle(abi.copy_config(cfg, precision=abi.LP_FP64))
    for n in TAVG_IN + (("PCon",) if npc else ()):
        o.set_field(n, lg.download(n).astype(np.float64))
    return o


@pytest.mark.parametrize("precision", [abi.LP_FP64, abi.LP_MIXED], ids=["fp64", "mixed"])
def test_energy_matches_oracle(precision):
    cfg, _, lg = _run(precision, 0)
    o = _oracle_of(cfg, lg, 0)
    kea_o, _ = o.energy()
    kea = lg.energy()
    print(f"kea gpu {kea:.16e} oracle {kea_o:.16e}")
    assert abs(kea - kea_o) <= 1e-14 * abs(kea_o)
    o.close()
    lg.close()


@pytest.mark.parametrize("precision,npc", [(abi.LP_FP64, 0), (abi.LP_FP64, 4), (abi.LP_MIXED, 4)],
                         ids=["fp64", "fp64-drops", "mixed-drops"])
def test_tavg_matches_oracle(precision, npc):
    cfg, ds, lg = _run(precision, npc)
    g = cfg.grid
    nq = 30 + 2 * npc if npc else 28
    o = _oracle_of(cfg, lg, npc)
    acc = np.zeros((nq, g.nz, g.ny, g.ld))
    o.tavg_compute(0.3, acc)
    lg.tavg_reset()
    lg.tavg_compute(0.3)
    # a second state: one more step on both sides
    lg.step(1)
    o.close()
    o = _oracle_of(cfg, lg, npc)
    o.tavg_compute(0.45, acc)
    lg.tavg_compute(0.45)
    worst = 0.0
    for q in range(nq):
        got, t = lg.tavg_get(q)
        assert t == pytest.approx(0.75, abs=1e-15)
        scale = max(np.abs(acc[q][..., :g.nx]).max(), 1e-300)
        err = np.abs(got[..., :g.nx] - acc[q][..., :g.nx]).max() / scale
        worst = max(worst, err)
        assert err <= TOL, (q, err)
        assert np.all(got[..., g.nx:] == 0.0)
    print(f"worst quantity error {worst:.2e} over {nq} accumulators")
    # tavg_init again: every accumulator and the total time back to zero
    lg.tavg_reset()
    z, t = lg.tavg_get(nq - 1)
    assert t == 0.0 and not np.any(z)
    o.close()
    lg.close()


def test_tavg_get_checks():
    from population_balance_amd.lesgo import LesgoError
    cfg, _, lg = _run(abi.LP_FP64, 0, nsteps=1)
    with pytest.raises(LesgoError):
        lg.tavg_get(0)  # nothing accumulated yet
    lg.tavg_reset()
    with pytest.raises(LesgoError):
        lg.tavg_get(28)  # no droplet quantities without droplet bins
    lg.close()
