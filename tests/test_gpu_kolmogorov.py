"""The 3-D viscous periodic Kolmogorov flow run as the configuration defines it: forward
with HBM snapshots, J = kinetic energy of the final state, and the discrete adjoint
started from dJ/dq(T) (the reference's nonzero adjoint terminal condition).  J and the
adjoint at t0 (dJ/dq0) against the oracle's same run (1e-12 / 1e-10), and dJ/dq0 against
central differences of J."""
import numpy as np
import pytest

import oracle_bind as ob
from magudi_amd import cases
from magudi_amd.functionals import kinetic_energy

pytestmark = pytest.mark.gpu


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("n", [16, 32])
def test_terminal_kinetic_energy_adjoint(n):
    from magudi_amd.solver import Region, Solver
    nsteps, K = 8, 3
    c = cases.kolmogorov3d(n=n, nsteps=nsteps)
    q0 = c.initial_state()
    R = Region(c)
    norm = R.grid(2)[0]
    S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=K)

    def cost(q):
        R.states[0].conservedVariables = q
        S.runForwardCheckpointed()
        return kinetic_energy(R.states[0].conservedVariables, norm, 3)

    J, lamT = cost(q0)
    S.runAdjointCheckpointed(terminal=lamT, want_gradient=False)
    lam0 = R.states[0].adjointVariables

    O = ob.Region(c)
    _, qT_o = O.run_forward(q0, c.dt, nsteps)
    J_o, lamT_o = kinetic_energy(qT_o, O.grid(2)[0], 3)
    _, _, lam0_o, _, _ = O.run_adjoint_ic(q0, c.dt, nsteps, K, lamT=lamT_o)
    assert abs(J - J_o) <= 1e-12 * abs(J_o)
    assert maxrel(lam0, lam0_o) <= 1e-10

    rng = np.random.default_rng(n)
    d = rng.standard_normal(q0.shape)
    d[0] *= 1e-2
    eps = 1e-6
    Jp, _ = cost(q0 + eps * d)
    Jm, _ = cost(q0 - eps * d)
    fd = (Jp - Jm) / (2 * eps)
    an = float(np.sum(norm * (lam0 * d)))
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize("n", [16, 32])
def test_device_terminal_kinetic_energy(n):
    """The terminal cost on the device (mgd_terminal_cost): J and dJ/dq(T) of the final
    state against the host formula on the same state, and the adjoint marched back from
    the device terminal state bit for bit equal to the one from the host terminal state."""
    from magudi_amd.solver import Region, Solver
    nsteps, K = 6, 4
    c = cases.kolmogorov3d(n=n, nsteps=nsteps)
    q0 = c.initial_state()
    R = Region(c, n_instances=2)
    norm = R.grid(2)[0]
    rng = np.random.default_rng(7)
    q1 = q0.copy()
    q1[1:4] += 0.01 * rng.standard_normal((3, c.npoints))
    R.states[0].conservedVariables = q0
    R.states[1].conservedVariables = q1
    S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=K)
    S.runForwardCheckpointed()
    qT = [R.states[k].conservedVariables for k in range(2)]
    J = R.terminalKineticEnergy()
    lam_dev = [R.states[k].adjointVariables for k in range(2)]
    S.runAdjointCheckpointed(terminal=True, want_gradient=False)
    lam0_dev = [R.states[k].adjointVariables for k in range(2)]
    host = [kinetic_energy(qT[k], norm, 3) for k in range(2)]
    S.runAdjointCheckpointed(terminal=np.stack([h[1] for h in host]), want_gradient=False)
    for k in range(2):
        Jh, lamh = host[k]
        assert abs(J[k] - Jh) <= 1e-13 * abs(Jh)
        assert np.array_equal(lam_dev[k], lamh)
        assert np.array_equal(R.states[k].adjointVariables, lam0_dev[k])


def test_instances_match_single_runs():
    """Two instances of a 32 x 16 x 16 periodic box in one region (every 3-D kernel,
    including the finishing kernels with the RK updates, indexes the instance) against
    each state run alone: bit for bit, forward and adjoint from a terminal state."""
    from magudi_amd.solver import Region, Solver
    from test_gpu_parity import periodic_box, perturbed
    c = periodic_box((32, 16, 16), dissipation=False)
    nsteps, K = 4, 2
    qs = [perturbed(c, 31), perturbed(c, 32)]
    rng = np.random.default_rng(33)
    lamT = [rng.standard_normal((c.nv, c.npoints)) for _ in range(2)]
    single = []
    for k in range(2):
        R = Region(c)
        R.states[0].conservedVariables = qs[k]
        S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=K)
        S.runForwardCheckpointed()
        qT = R.states[0].conservedVariables
        S.runAdjointCheckpointed(terminal=lamT[k], want_gradient=False)
        single.append((qT, R.states[0].adjointVariables))
    R = Region(c, n_instances=2)
    for k in range(2):
        R.states[k].conservedVariables = qs[k]
    S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=K)
    S.runForwardCheckpointed()
    qT2 = [R.states[k].conservedVariables for k in range(2)]
    S.runAdjointCheckpointed(terminal=np.stack(lamT), want_gradient=False)
    for k in range(2):
        assert np.array_equal(qT2[k], single[k][0]), k
        assert np.array_equal(R.states[k].adjointVariables, single[k][1]), k
