"""Tile shapes and the autotuner of the fused 2-D stage kernels.

No result may depend on the tile shape: every rhs value is computed by the same
arithmetic whatever tile holds the point, and J / sensitivity are reduced from
per-point terms in a fixed order.  So the shapes must agree bit for bit, and each
must match the oracle to the parity tolerances (rhs 1e-13, J 1e-12, gradient 1e-10).
"""
import numpy as np
import pytest

import oracle_bind as ob
from magudi_amd import cases

pytestmark = pytest.mark.gpu

SHAPES = [0, 1, 2, 3, 4, 5]


def maxrel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def periodic_x_monopole(n=48):
    """Monopole with x periodic (PLANE): far field and sponges on the y faces only."""
    c = cases.acoustic_monopole(n=n, nsteps=1)
    dx = 28.0 / (n - 1)
    c.periodicity = (cases.PLANE, cases.NONE, cases.NONE)
    c.periodic_length = (n * dx, 0.0, 0.0)
    c.patches = [p for p in c.patches if p.normal_dir not in (1, -1)]
    return c


CASES = {
    "monopole60": lambda: cases.acoustic_monopole(n=60, nsteps=1),
    "monopole60_composite": lambda: cases.acoustic_monopole(n=60, nsteps=1, composite=True),
    "periodic_x48": periodic_x_monopole,
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("mode", [1, -1])
def test_rhs_all_shapes(name, mode):
    from magudi_amd.solver import Region
    c = CASES[name]()
    O = ob.Region(c)
    rng = np.random.default_rng(31)
    q = c.target + 0.05 * rng.standard_normal((c.nv, c.npoints))
    q[0] = np.abs(q[0]) + 0.2
    lam = rng.standard_normal((c.nv, c.npoints))
    outs = []
    for sh in SHAPES:
        R = Region(c, tile_shape=sh)
        assert R.tile_shape == sh
        R.states[0].conservedVariables = q
        R.states[0].adjointVariables = lam
        na = R.n_actuator
        ctl = np.linspace(-1.0, 1.0, na) if na else None
        if mode == 1:
            R.computeRhs(1, time=2.3, control=ctl)
        else:
            R.updateAdjointForcing()
            R.computeRhs(-1, forcing_factor=1.0)
        outs.append(R.states[0].rightHandSide.copy())
    if mode == 1:
        ref = O.rhs(1, q, time=2.3, control=np.linspace(-1.0, 1.0, R.n_actuator) if R.n_actuator else None)
    else:
        af = O.adjoint_forcing(q) if ob.lib().orc_cost_target_npoints(O.h) > 0 else None
        ref = O.rhs(-1, q, lam, aforcing=af, factor=1.0)
    for sh, o in zip(SHAPES, outs):
        assert np.array_equal(o, outs[0]), sh
        assert maxrel(o, ref) < 1e-13, (sh, maxrel(o, ref))


@pytest.mark.parametrize("name", ["monopole60", "periodic_x48"])
def test_adjoint_run_all_shapes(name):
    from magudi_amd.solver import Region, Solver
    c = CASES[name]()
    c.nsteps, c.save_interval, c.dt = 10, 4, 0.2
    q0 = c.initial_state()
    res = []
    for sh in SHAPES + [-1]:
        R = Region(c, tile_shape=sh)
        R.states[0].conservedVariables = q0
        sens, grad, J = Solver(R, dt=c.dt, nsteps=c.nsteps, save_interval=c.save_interval).runAdjoint()
        assert R.tile_shape in SHAPES
        res.append((sens[0], grad[0].copy(), J[0]))
    for sh, (s, g, j) in zip(SHAPES + [-1], res):
        assert s == res[0][0] and j == res[0][2] and np.array_equal(g, res[0][1]), sh
    so, go, _, Jo = ob.Region(c).run_adjoint(q0, c.dt, c.nsteps, c.save_interval)
    assert abs(res[0][2] - Jo) <= 1e-12 * abs(Jo)
    assert abs(res[0][0] - so) <= 1e-10 * abs(so)
    assert np.abs(res[0][1] - go).max() <= 1e-10 * np.abs(go).max()
