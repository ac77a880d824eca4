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


@pytest.mark.parametrize("nsteps", [3, 4, 9, 12])
@pytest.mark.parametrize("graph", [False, True])
def test_window_counts_match_resident(nsteps, graph):
    """The forward pass hands the stage states of its last window to the adjoint pass;
    one window (nsteps <= K), a ragged last window and whole windows, eager and in a
    captured graph, against the resident states bit for bit."""
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=41, nsteps=nsteps, save_interval=4)
    c.dt = 0.2
    q0 = c.initial_state()
    out = []
    for resident in (True, False):
        R = Region(c, resident_states=resident)
        if not resident:
            R.set_step_store(0)
        ctl = 1e-2 * np.random.default_rng(nsteps).standard_normal((4 * nsteps, R.n_actuator))
        S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=c.save_interval)
        for rep in range(2 if graph else 1):
            R.states[0].conservedVariables = q0
            sens, grad, J = S.runAdjoint(control=ctl, use_graph=graph)
            res = (sens.copy(), grad.copy(), J.copy(), R.states[0].adjointVariables)
            if rep:
                assert all(np.array_equal(a, b) for a, b in zip(res, out[-1])), "graph replay"
            else:
                out.append(res)
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("ninst", [1, 3])
def test_fused_step_store_matches_resident(ninst):
    """The fused 2-D path with the step store (MGD_OPT_STEP_STORE): the state at the start
    of every step kept, the inner stage states of none, of the last save_interval or of
    every step held, the others recomputed three stages per step in the launches of the
    adjoint stages -- against the resident stage states and the window recompute, bit for
    bit, eager and as two graph replays."""
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=41, nsteps=9, save_interval=4)
    c.dt = 0.2
    q0 = c.initial_state()
    plans = [("resident", None, None, {"mode": "fused", "held": 37, "first": 0}),
             ("windows", 0, 0, {"mode": "fused", "held": 0, "first": 0}),
             ("step store", 1, 0, {"mode": "step store", "held": 0, "first": 9}),
             ("step store", 1, 1, {"mode": "step store", "held": 4, "first": 5}),
             ("step store", 1, -1, {"mode": "step store", "held": 9, "first": 0})]
    ref = None
    for name, step, cache, want in plans:
        print(f"fused plan {name} step={step} cache={cache} ninst={ninst}", flush=True)
        R = Region(c, n_instances=ninst, resident_states=name == "resident")
        if step is not None:
            R.set_step_store(step)
            R.set_window_cache(cache)
        na = R.n_actuator
        rng = np.random.default_rng(19)
        ctl = np.stack([1e-2 * (k + 1) * rng.standard_normal((4 * c.nsteps, na)) for k in range(ninst)])
        S = Solver(R, dt=c.dt, nsteps=c.nsteps, save_interval=c.save_interval)
        for graph in (False, True, True):
            for i in range(ninst):
                R.states[i].conservedVariables = q0
            sens, grad, J = S.runAdjoint(control=ctl, use_graph=graph)
            lam = np.stack([R.states[i].adjointVariables for i in range(ninst)])
            res = (sens.copy(), grad.copy(), J.copy(), lam)
            assert R.checkpoint_plan == want, (name, R.checkpoint_plan)
            if ref is None:
                ref = res
                assert np.abs(lam).max() > 0 and np.all(np.abs(sens) > 0)
            for a, b in zip(res, ref):
                assert np.array_equal(a, b), (name, step, cache, graph)
        R.close()


def test_split_passes_and_repeated_adjoint():
    """Forward and adjoint passes called one by one (the multiple-shooting path): the
    first adjoint pass takes the last window from the forward pass, a second adjoint
    pass from the same snapshots recomputes it; both equal the one-call run."""
    from magudi_amd import capi
    from magudi_amd.capi import dptr
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=41, nsteps=10, save_interval=4)
    c.dt = 0.2
    q0 = c.initial_state()
    R = Region(c)
    na = R.n_actuator
    ctl = 1e-2 * np.random.default_rng(5).standard_normal((4 * c.nsteps, na))
    R.states[0].conservedVariables = q0
    s_ref, g_ref, J_ref = Solver(R, dt=c.dt, nsteps=c.nsteps, save_interval=4).runAdjoint(control=ctl)
    lam_ref = R.states[0].adjointVariables
    R.states[0].conservedVariables = q0
    J = np.zeros(1)
    capi.check(capi.lib().mgd_run_forward_checkpointed(R.handle, 0.0, c.dt, c.nsteps, 4,
                                                       dptr(np.ascontiguousarray(ctl)), dptr(J)))
    assert J[0] == J_ref[0]
    for rep in range(2):
        grad = np.zeros((4 * c.nsteps, na))
        sens = np.zeros(1)
        capi.check(capi.lib().mgd_run_adjoint_checkpointed(R.handle, 0, dptr(grad), dptr(sens)))
        assert sens[0] == s_ref[0], rep
        assert np.array_equal(grad, g_ref.reshape(grad.shape)), rep
        assert np.array_equal(R.states[0].adjointVariables, lam_ref), rep
