"""Region%getTimeStepSize in constant-CFL mode and State%computeCfl on the device
(computeTimeStepSize / computeCfl, src/CNSHelperImpl.f90:842-982) against the oracle on the
same perturbed states: min / max reductions are exact, and the per-point arithmetic is the
reference's (the off-diagonal metrics of a rectilinear grid only add exact zeros), so the
values agree to the last bit."""
import numpy as np
import pytest

import oracle_bind as ob
from magudi_amd import cases

pytestmark = pytest.mark.gpu


def _perturbed(c, seed, amp):
    rng = np.random.default_rng(seed)
    q = c.initial_state() + amp * rng.standard_normal((c.nv, c.npoints))
    q[0] = np.abs(q[0]) + 0.3
    return q


CASES = {
    "monopole41": lambda: cases.acoustic_monopole(n=41, nsteps=1),
    "walled41_viscous": lambda: cases.walled_monopole(n=41, viscous=True),
    "kolmogorov16": lambda: cases.kolmogorov3d(n=16, nsteps=1),
    "wave1d": lambda: cases.periodic_wave_1d(n=256, nsteps=1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_time_step_and_cfl_match_oracle(name):
    from magudi_amd.solver import Region
    c = CASES[name]()
    R = Region(c, n_instances=2)
    O = ob.Region(c)
    qs = [_perturbed(c, 1, 0.02), _perturbed(c, 2, 0.1)]
    for k, q in enumerate(qs):
        R.states[k].conservedVariables = q
    cfl = 0.5
    dts = R.computeTimeStepSize(cfl)
    cfls = R.computeCfl(1.3e-3)
    for k, q in enumerate(qs):
        assert dts[k] == O.time_step_size(q, cfl), (k, dts[k], O.time_step_size(q, cfl))
        assert cfls[k] == O.cfl(q, 1.3e-3)
    assert R.getTimeStepSize(cfl) == dts.min()
    assert dts[0] != dts[1]
