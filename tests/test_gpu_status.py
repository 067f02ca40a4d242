"""Failure status of the step: a non-finite velocity is reported by lp_step
(LP_ERR_NAN) and by lp_check_finite with the field's name; an optional CFL limit makes
lp_step return LP_ERR_CFL (the reference only reports the CFL, main.f90:902-926).  The
conditions are planted in the inputs from the host; every kernel runs normally."""
import numpy as np
import pytest

from harness import seeded_state
from population_balance_amd import abi

pytestmark = pytest.mark.gpu


def _ctx(prec):
    from population_balance_amd.lesgo import Lesgo
    cfg = abi.make_config(32, 32, 17, precision=prec)
    lg = Lesgo(cfg)
    for k, a in seeded_state(cfg).items():
        lg.upload(k, a)
    return cfg, lg


@pytest.mark.parametrize("prec", [abi.LP_FP64, abi.LP_MIXED])
def test_nonfinite_velocity_reported(prec):
    from population_balance_amd.lesgo import LesgoError
    cfg, lg = _ctx(prec)
    lg.check_finite()  # a finite state passes
    lg.step(1)
    u = lg.download("u")
    u[3, 5, 7] = np.nan
    lg.upload("u", u)
    with pytest.raises(LesgoError, match="u holds") as e:
        lg.check_finite()
    assert e.value.code == abi.LP_ERR_NAN
    with pytest.raises(LesgoError, match="non-finite velocity") as e:
        lg.step(1)
    assert e.value.code == abi.LP_ERR_NAN
    assert lg.time()[0] == 3  # the step ran; the flag is cleared for the next call
    lg.close()


@pytest.mark.parametrize("prec", [abi.LP_FP64, abi.LP_MIXED])
def test_cfl_limit(prec):
    from population_balance_amd.lesgo import LesgoError
    cfg, lg = _ctx(prec)
    lg.step(1)
    cfl, _ = lg.timestep_conditions()
    lg.set_cfl_limit(2.0 * cfl)
    lg.step(1)  # well below the limit
    lg.set_cfl_limit(0.5 * cfl)
    with pytest.raises(LesgoError, match="CFL") as e:
        lg.step(1)
    assert e.value.code == abi.LP_ERR_CFL
    lg.set_cfl_limit(0.0)  # off again
    lg.step(1)
    assert lg.precision() == prec
    lg.close()
