"""The monopole configurations at their full grid sizes.

C2's own gradient check (401^2, 960 RK4 steps, snapshots every 32): the finite-difference
relation (J(A) - J(0)) / A at A = 1e-4 along the adjoint gradient against Grad0 (the
check README.md:80-157 of the reference describes).  Against the oracle at full grid
size (the oracle's per-point loops on the host's threads): C2's 401^2 forward + adjoint
over 320 steps (t = 7.9, the pulse inside the cost annulus; K = 32) for J, Grad0 and the
gradient vector, and C1's 201^2 forward run of
480 steps for J."""
import numpy as np
import pytest

from magudi_amd import cases

pytestmark = pytest.mark.gpu


def test_monopole401_gradient_fd():
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=401, nsteps=960, save_interval=32)
    q0 = c.initial_state()
    R = Region(c)
    R.states[0].conservedVariables = q0
    S = Solver(R)
    sens, grad, J = S.runAdjoint()
    assert np.isfinite(J[0]) and sens[0] > 0.0
    R.states[0].conservedVariables = q0
    A = 1e-4
    J1 = S.runForward(control=A * grad[0])[0]
    fd = (J1 - J[0]) / A
    assert abs(fd - sens[0]) <= 1e-3 * abs(sens[0]), (fd, sens[0])


def test_monopole401_against_oracle_320_steps():
    import oracle_bind as ob
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=401, nsteps=320, save_interval=32)
    q0 = c.initial_state()
    R = Region(c)
    R.states[0].conservedVariables = q0
    sens, grad, J = Solver(R).runAdjoint()
    so, go, lo, Jo = ob.Region(c).run_adjoint(q0, c.dt, c.nsteps, c.save_interval)
    assert abs(J[0] - Jo) <= 1e-12 * abs(Jo)
    assert abs(sens[0] - so) <= 1e-10 * abs(so)
    assert np.abs(grad[0] - go).max() <= 1e-10 * np.abs(go).max()
    lam = R.states[0].adjointVariables
    assert np.abs(lam - lo).max() <= 1e-10 * np.abs(lo).max()


def test_monopole201_forward_against_oracle_480_steps():
    import oracle_bind as ob
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=201, nsteps=480, save_interval=32)
    q0 = c.initial_state()
    R = Region(c)
    R.states[0].conservedVariables = q0
    J = Solver(R).runForward()[0]
    Jo, qo = ob.Region(c).run_forward(q0, c.dt, c.nsteps)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    q = R.states[0].conservedVariables
    assert np.abs((q - q0) - (qo - q0)).max() <= 1e-11 * np.abs(qo - q0).max()
