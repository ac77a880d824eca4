"""Solution limits (checkSolutionLimits, src/SolverImpl.f90:189-304, 814-819): with wide
limits a run is unchanged bit for bit; an instance whose temperature reaches the bound
gets J = HUGE (and sensitivity = HUGE in adjoint runs) and the run reports it, while the
other instances of the batch keep their values; the host-loop check flags a state with a
density at the bound or a NaN."""
import sys

import numpy as np
import pytest

from magudi_amd import cases

pytestmark = pytest.mark.gpu

HUGE = sys.float_info.max


def _case():
    c = cases.acoustic_monopole(n=41, nsteps=10, save_interval=4)
    c.dt = 0.2
    return c


@pytest.mark.parametrize("plain", [False, True])
def test_limits_flag_only_the_violating_instance(plain):
    from magudi_amd.solver import Region, Solver
    c = _case()
    q0 = c.initial_state()
    # instance 1 starts warmer: its temperature is above the bound at the first stage
    q1 = q0.copy()
    q1[3] = q1[3] * 1.1
    ref = []
    for q in (q0, q1):
        R = Region(c, force_plain=plain)
        R.states[0].conservedVariables = q
        sens, _, J = Solver(R).runAdjoint()
        ref.append((J[0], sens[0]))
        assert not R.solutionCrashed
    R = Region(c, n_instances=2, force_plain=plain)
    R.setSolutionLimits((0.5, 1.5), (2.0, 2.6))    # quiescent temperature is 2.5
    for k, q in enumerate((q0, q1)):
        R.states[k].conservedVariables = q
    S = Solver(R)
    sens, _, J = S.runAdjoint()
    assert R.solutionCrashed
    assert J[0] == ref[0][0] and sens[0] == ref[0][1]
    assert J[1] == HUGE and sens[1] == HUGE
    # forward runs report the same
    for k, q in enumerate((q0, q1)):
        R.states[k].conservedVariables = q
    J = S.runForward()
    assert R.solutionCrashed and J[0] == ref[0][0] and J[1] == HUGE
    # wide limits: nothing flagged, results unchanged
    R.setSolutionLimits((0.1, 10.0), (0.1, 10.0))
    for k, q in enumerate((q0, q1)):
        R.states[k].conservedVariables = q
    sens, _, J = S.runAdjoint()
    assert not R.solutionCrashed
    assert J[0] == ref[0][0] and J[1] == ref[1][0] and sens[1] == ref[1][1]


def test_host_loop_check():
    from magudi_amd.solver import Region
    c = _case()
    R = Region(c, n_instances=3)
    R.setSolutionLimits((0.5, 1.5), (2.0, 3.0))
    q = c.initial_state()
    qa = q.copy()
    qa[0, 100] = 0.5            # density at the minimum: out of range (f <= min)
    qn = q.copy()
    qn[3, 7] = np.nan
    for k, s in enumerate((q, qa, qn)):
        R.states[k].conservedVariables = s
    assert list(R.checkSolutionLimits()) == [False, True, True]
