"""Resident stage states (MGD_OPT_RESIDENT_STATES): the forward run keeps every RK
stage state in HBM and the adjoint reads them instead of recomputing windows from the
snapshots.  The states are the same arrays either way, so J, the sensitivity, the
gradient and the adjoint state must agree bit for bit with the default
checkpoint / recompute path, and with the oracle to the parity tolerances."""
import numpy as np
import pytest

import oracle_bind as ob
from magudi_amd import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ninst", [1, 3])
def test_resident_matches_recompute(ninst):
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=41, nsteps=14, save_interval=4)
    c.dt = 0.2
    q0 = c.initial_state()
    rng = np.random.default_rng(9)
    out = []
    for resident in (False, True):
        R = Region(c, n_instances=ninst, resident_states=resident)
        for i in range(ninst):
            R.states[i].conservedVariables = q0
        na = R.n_actuator
        ctl = np.stack([1e-2 * k * rng.standard_normal((4 * c.nsteps, na)) for k in range(ninst)])
        rng = np.random.default_rng(9)
        sens, grad, J = Solver(R, dt=c.dt, nsteps=c.nsteps, save_interval=c.save_interval).runAdjoint(control=ctl)
        lam = np.stack([R.states[i].adjointVariables for i in range(ninst)])
        out.append((sens.copy(), grad.copy(), J.copy(), lam))
    (s0, g0, j0, l0), (s1, g1, j1, l1) = out
    assert np.array_equal(j0, j1) and np.array_equal(s0, s1)
    assert np.array_equal(g0, g1) and np.array_equal(l0, l1)
    if ninst == 1:
        so, go, lo, Jo = ob.Region(c).run_adjoint(q0, c.dt, c.nsteps, c.save_interval)
        assert abs(j1[0] - Jo) <= 1e-12 * abs(Jo)
        assert abs(s1[0] - so) <= 1e-10 * abs(so)
        assert np.abs(g1[0] - go).max() <= 1e-10 * np.abs(go).max()
