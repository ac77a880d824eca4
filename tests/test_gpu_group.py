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


@pytest.mark.parametrize("resident", [True, False])
def test_merged_group_equals_single_runs(resident):
    """One launch per RK stage for all fused 2-D members (mgd_group_run_adjoint), the 1-D
    member beside them: J, sensitivity and gradient of every member equal its own run
    bit for bit, with a control on one member and two instances on another, the members'
    stage states resident (MGD_OPT_RESIDENT_STATES 1 on every member) or on snapshots and
    windows (the default); the members' tile shapes are back to their own after the run."""
    from magudi_amd import capi
    from magudi_amd.group import RegionGroup
    from magudi_amd.solver import Region, Solver
    cs = small_batch()
    ninst = [1, 1, 2, 2]
    G = RegionGroup(cs, n_instances=ninst)
    shapes = [r.tile_shape for r in G.regions]
    assert G._merged_members() == [1, 2, 3]
    if resident:
        for k in (1, 2, 3):
            capi.check(capi.lib().mgd_region_set_option(G.regions[k].handle, capi.OPT_RESIDENT_STATES, 1))
    rng = np.random.default_rng(17)
    na = G.regions[2].n_actuator
    controls = [None, None, 1e-2 * rng.standard_normal((2, 4 * cs[2].nsteps, na)), None]
    for rep in range(2):   # second run reuses the group and its tables
        G.restart()
        out = G.run_adjoint_merged(controls=controls)
        for k in (1, 2, 3):
            plan = G.regions[k].checkpoint_plan
            assert plan == {"mode": "fused", "held": 4 * cs[k].nsteps + 1 if resident else 0,
                            "first": 0}, (k, plan)
        for k, c in enumerate(cs):
            R = Region(c, n_instances=ninst[k])
            for i in range(R.n_instances):
                R.states[i].conservedVariables = c.initial_state()
            s1, g1, J1 = Solver(R).runAdjoint(control=controls[k])
            s, g, J = out[k]
            assert np.array_equal(J, J1) and np.array_equal(s, s1), (rep, c.name)
            assert (g is None and g1 is None) or np.array_equal(g, g1), (rep, c.name)
    assert [r.tile_shape for r in G.regions] == shapes
    if resident:
        # resident on some members but not all: refused
        capi.check(capi.lib().mgd_region_set_option(G.regions[2].handle, capi.OPT_RESIDENT_STATES, -1))
        with pytest.raises(capi.MagudiError):
            G.run_adjoint_merged(controls=controls)
    G.close()


def shooting_inputs(O, c, nsplit, nts, seed=11):
    """Segment initial conditions of a mixing layer as msforward reads them: the states of
    one oracle run at t = k nts dt, each but the first with a seeded momentum bump (so the
    matching conditions are not met), and time offsets k nts dt."""
    x, y = c.coords[0], c.coords[1]
    rng = np.random.default_rng(seed)
    q = c.initial_state()
    ics, t0s = [q], [0.0]
    for k in range(1, nsplit):
        _, q = O.run_forward(q, c.dt, nts, t0=(k - 1) * nts * c.dt)
        bump = np.exp(-((x - rng.uniform(5, 25)) ** 2 + (y - rng.uniform(-2, 2)) ** 2) / 4.0)
        ic = q.copy()
        ic[1] = ic[1] + 1e-3 * bump
        ic[2] = ic[2] - 5e-4 * bump
        ics.append(ic)
        t0s.append(k * nts * c.dt)
    return ics, t0s


@pytest.mark.parametrize("merged", [False, True])
def test_group_shooting_member_matches_oracle(merged):
    """The mixing layer as a multiple-shooting member of an uneven group (4 segments with
    their own initial conditions and time offsets, quadratic penalty, a control on the
    segments): J, the segment costs and mismatches, the control and IC gradients against
    the oracle's msforward / msadjoint (J 1e-12, gradients 1e-10); the plain members and
    the shooting member itself bit for bit against their runs alone."""
    import ms_oracle as ms
    from magudi_amd.group import RegionGroup
    from magudi_amd.multishoot import MultipleShooting
    from magudi_amd.solver import Region, Solver
    nsplit, nts, K, w = 4, 4, 2, 20.0
    cs = [cases.periodic_wave_1d(n=256, nsteps=6, save_interval=3),
          cases.acoustic_monopole(n=45, nsteps=6, save_interval=3),
          cases.mixing_layer(nx=64, ny=45, nsteps=nts, save_interval=K)]
    cs[1].dt = 0.2
    c = cs[2]
    O = ob.Region(c)
    ics, t0s = shooting_inputs(O, c, nsplit, nts)
    na = O.n_actuator
    assert na > 0
    ctl = 1e-2 * np.random.default_rng(5).standard_normal((nsplit * 4 * nts, na))
    opt = dict(nsplit=nsplit, nts=nts, penalty_weight=w, save_interval=K, ics=ics, t0s=t0s)
    G = RegionGroup(cs, shooting={2: opt})
    assert G.regions[2].n_instances == nsplit
    if merged:
        assert 2 not in G._merged_members()
    G.restart()
    controls = [None, None, ctl]
    out = G.run_adjoint_merged(controls=controls) if merged else G.run_adjoint(controls=controls)
    Jo, cost_o, L2_o, _ = ms.ms_forward(O, ics, t0s, c.dt, nts, w, control=ctl)
    res_o = ms.ms_adjoint(O, ics, t0s, c.dt, nts, K, w, control=ctl)
    total, grad, J = out[2]
    M = G.shooting[2][0]
    assert abs(J[0] - Jo) <= 1e-12 * abs(Jo)
    assert rel(M.segment_cost, cost_o) < 1e-12
    assert min(L2_o[1:]) > 0 and rel(M.L2sq[1:], L2_o[1:]) < 1e-12
    assert abs(total[0] - res_o["total"]) <= 1e-10 * abs(res_o["total"])
    assert rel(grad, res_o["control_grad"]) < 1e-10
    res = G.shooting_results[2]
    for k in range(nsplit):
        assert rel(res["ic_grad"][k], res_o["ic_grad"][k]) < 1e-10, k
    # the shooting member alone, and the plain members alone: bit for bit
    A = MultipleShooting(c, nsplit=nsplit, nts=nts, penalty_weight=w, save_interval=K)
    assert A.msforward(ics, t0s, control=ctl) == J[0]
    a = A.msadjoint()
    assert a["total"] == total[0] and np.array_equal(a["control_grad"], grad)
    for k in range(2):
        R = Region(cs[k])
        R.states[0].conservedVariables = cs[k].initial_state()
        s1, g1, J1 = Solver(R).runAdjoint()
        s, g, Jk = out[k]
        assert np.array_equal(Jk, J1) and np.array_equal(s, s1), cs[k].name
    G.close()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
