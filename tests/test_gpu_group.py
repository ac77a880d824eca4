"""Uneven batches (magudi_amd/group.py): 1-D waves, monopoles of different sizes and a
mixing layer run concurrently must give exactly what each region gives alone, and
agree with the generic per-instance path and with the oracle (J 1e-12, gradients
1e-10)."""
import numpy as np
import pytest

import oracle_bind as ob
from magudi_amd import cases

pytestmark = pytest.mark.gpu


def small_batch():
    cs = [cases.periodic_wave_1d(n=256, nsteps=6, save_interval=3),
          cases.acoustic_monopole(n=45, nsteps=6, save_interval=3),
          cases.acoustic_monopole(n=61, nsteps=6, save_interval=3),
          cases.mixing_layer(nx=64, ny=45, nsteps=6, save_interval=3)]
    for c in cs:
        if c.nd == 2 and c.name.startswith("monopole"):
            c.dt = 0.2
    return cs


def test_group_equals_single_runs_and_plain_path():
    from magudi_amd.group import RegionGroup
    from magudi_amd.solver import Region, Solver
    cs = small_batch()
    G = RegionGroup(cs, n_instances=[1, 1, 1, 2])
    G.restart()
    out = G.run_adjoint()
    for k, c in enumerate(cs):
        R = Region(c, n_instances=G.regions[k].n_instances)
        for i in range(R.n_instances):
            R.states[i].conservedVariables = c.initial_state()
        s1, g1, J1 = Solver(R).runAdjoint()
        s, g, J = out[k]
        assert np.array_equal(J, J1) and np.array_equal(s, s1), c.name
        assert np.array_equal(g, g1), c.name
        P = Region(c, force_plain=True)
        P.states[0].conservedVariables = c.initial_state()
        sp, gp, Jp = Solver(P).runAdjoint()
        assert abs(J[0] - Jp[0]) <= 1e-12 * abs(Jp[0]), c.name
        assert abs(s[0] - sp[0]) <= 1e-10 * abs(sp[0]) + 1e-300, c.name
        assert np.abs(g[0] - gp[0]).max() <= 1e-10 * np.abs(gp[0]).max() + 1e-300, c.name


@pytest.mark.parametrize("which", [0, 3])
def test_group_member_matches_oracle(which):
    from magudi_amd.group import RegionGroup
    cs = small_batch()
    G = RegionGroup(cs)
    G.restart()
    s, g, J = G.run_adjoint()[which]
    c = cs[which]
    so, go, _, Jo = ob.Region(c).run_adjoint(c.initial_state(), c.dt, c.nsteps, c.save_interval)
    assert abs(J[0] - Jo) <= 1e-12 * abs(Jo)
    assert abs(s[0] - so) <= 1e-10 * abs(so) + 1e-300
    assert np.abs(g[0] - go).max() <= 1e-10 * np.abs(go).max() + 1e-300


def test_merged_group_equals_single_runs():
    """One launch per RK stage for all fused 2-D members (mgd_group_run_adjoint), the 1-D
    member beside them: J, sensitivity and gradient of every member equal its own run
    bit for bit, with a control on one member and two instances on another; the members'
    tile shapes are back to their own after the run."""
    from magudi_amd.group import RegionGroup
    from magudi_amd.solver import Region, Solver
    cs = small_batch()
    ninst = [1, 1, 2, 2]
    G = RegionGroup(cs, n_instances=ninst)
    shapes = [r.tile_shape for r in G.regions]
    assert G._merged_members() == [1, 2, 3]
    rng = np.random.default_rng(17)
    na = G.regions[2].n_actuator
    controls = [None, None, 1e-2 * rng.standard_normal((2, 4 * cs[2].nsteps, na)), None]
    for rep in range(2):   # second run reuses the group and its tables
        G.restart()
        out = G.run_adjoint_merged(controls=controls)
        for k, c in enumerate(cs):
            R = Region(c, n_instances=ninst[k])
            for i in range(R.n_instances):
                R.states[i].conservedVariables = c.initial_state()
            s1, g1, J1 = Solver(R).runAdjoint(control=controls[k])
            s, g, J = out[k]
            assert np.array_equal(J, J1) and np.array_equal(s, s1), (rep, c.name)
            assert (g is None and g1 is None) or np.array_equal(g, g1), (rep, c.name)
    assert [r.tile_shape for r in G.regions] == shapes
    G.close()
