"""A run driven stage by stage through the C ABI, as the reference's Solver loops drive it
(src/SolverImpl.f90:672-912 forward, 914-1245 adjoint): RK4Integrator.substepForward /
substepAdjoint, UniformCheckpointer.migrateTo, the functional after every forward stage,
the control gradient and sensitivity before every adjoint stage, the adjoint forcing
switched off at the final stage.  The states must equal mgd_run_forward /
mgd_run_adjoint_checkpointed's bit for bit (same kernels in the same order); J and the
sensitivity, summed in another order than the runs' fixed-tree quadrature, to 1e-12."""
import numpy as np
import pytest

from magudi_amd import cases

pytestmark = pytest.mark.gpu


def _case(walled=False):
    if walled:
        c = cases.walled_monopole(n=41, viscous=True, nsteps=7, save_interval=3)
    else:
        c = cases.acoustic_monopole(n=41, nsteps=11, save_interval=4)
    c.dt = 0.2
    return c


@pytest.mark.parametrize("plain,walled", [(False, False), (True, False), (True, True)])
def test_stagewise_run_matches_whole_run(plain, walled):
    from magudi_amd.solver import Region, RK4Integrator, Solver, UniformCheckpointer
    c = _case(walled)
    print(f"stage api: {c.name} plain={plain}", flush=True)
    q0 = c.initial_state()
    R = Region(c, n_instances=2, force_plain=plain)
    na = R.n_actuator
    rng = np.random.default_rng(3)
    ctl = None
    if na:
        ctl = np.stack([a * rng.standard_normal((4 * c.nsteps, na)) for a in (0.0, 2e-2)])
    for k in range(2):
        R.states[k].conservedVariables = q0
    S = Solver(R, dt=c.dt, nsteps=c.nsteps, save_interval=c.save_interval)
    J_run = S.runForwardCheckpointed(control=ctl)
    qT_run = [R.states[k].conservedVariables for k in range(2)]
    sens_run, grad_run = S.runAdjointCheckpointed()
    lam0_run = [R.states[k].adjointVariables for k in range(2)]

    # forward, stage by stage (Solver%runForward)
    ti = RK4Integrator()
    for k in range(2):
        R.states[k].conservedVariables = q0
    J = np.zeros(2)
    dt, time = c.dt, 0.0
    for step in range(1, c.nsteps + 1):
        for i in range(1, 5):
            if i in (2, 4):          # State%time as substepForwardRK4 advances it
                time = time + dt / 2.0
            sidx = 4 * (step - 1) + i - 1
            ti.substepForward(R, time, dt, step, i, None if ctl is None else ctl[:, sidx])
            J += ti.norm[i - 1] * dt * R.functional()
    for k in range(2):
        assert np.array_equal(R.states[k].conservedVariables, qT_run[k]), k
    assert np.allclose(J, J_run, rtol=1e-12, atol=0.0)

    # adjoint, stage by stage over the snapshots of the checkpointed run (from q0 again:
    # the stage-wise forward left the final state in the conserved variables)
    for k in range(2):
        R.states[k].conservedVariables = q0
    S.runForwardCheckpointed(control=ctl)
    R.adjointInitialCondition()
    mig = UniformCheckpointer()
    sens = np.zeros(2)
    grad = np.zeros((2, 4 * c.nsteps, na)) if na else None
    for step in range(c.nsteps - 1, -1, -1):
        for i in range(4, 0, -1):
            if i == 1:
                mig.migrateTo(R, step, 4)
            else:
                mig.migrateTo(R, step + 1, i - 1)
            g, s = R.controlGradient()
            if na:
                grad[:, 4 * step + i - 1] = g
            sens += ti.norm[i - 1] * dt * s
            R.updateAdjointForcing(no_adjoint_forcing=(step == 0 and i == 1))
            ti.substepAdjoint(R, 0.0, dt, step, i)
    for k in range(2):
        assert np.array_equal(R.states[k].adjointVariables, lam0_run[k]), k
    if na:
        assert np.array_equal(grad, grad_run)
        assert np.allclose(sens, sens_run, rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("step,cache,fused", [(-1, -1, False), (0, -1, False), (0, 0, False),
                                              (1, 0, False), (1, 1, False), (1, 0, True),
                                              (1, 1, True), (0, 0, True)])
def test_migrate_serves_every_stage_state(step, cache, fused):
    """migrateTo in any order, across windows, against the stage states of a run that keeps
    every stage (MGD_OPT_RESIDENT_STATES is not needed: a whole forward run with the
    stage outputs collected by substeps); with the step store (MGD_OPT_STEP_STORE, every
    step's inner states held, none, or those of the last save_interval steps) and with
    snapshots and windows (every window cached, none), on the generic kernels and on the
    fused 2-D stage kernels (stage states not resident)."""
    from magudi_amd.solver import Region, RK4Integrator, Solver, UniformCheckpointer
    c = _case()
    q0 = c.initial_state()
    R = Region(c, force_plain=True) if not fused else Region(c, resident_states=False)
    assert (R.tile_shape >= 0) == fused
    R.set_step_store(step)
    R.set_window_cache(cache)
    R.states[0].conservedVariables = q0
    ti = RK4Integrator()
    states = {(0, 4): q0}
    time = 0.0
    for ts in range(1, c.nsteps + 1):
        for i in range(1, 5):
            if i in (2, 4):
                time = time + c.dt / 2.0
            ti.substepForward(R, time, c.dt, ts, i)
            states[(ts, i)] = R.states[0].conservedVariables
    R.states[0].conservedVariables = q0
    S = Solver(R, dt=c.dt, nsteps=c.nsteps, save_interval=c.save_interval)
    S.runForwardCheckpointed()
    plan = R.checkpoint_plan
    if step == 1:
        assert plan == {"mode": "step store", "held": 0 if cache == 0 else c.save_interval,
                        "first": c.nsteps - (0 if cache == 0 else c.save_interval)}, plan
    elif fused:
        assert plan == {"mode": "fused", "held": 0, "first": 0}, plan
    mig = UniformCheckpointer()
    order = list(states)
    np.random.default_rng(9).shuffle(order)
    for key in order + sorted(states, reverse=True):
        mig.migrateTo(R, *key)
        assert np.array_equal(R.states[0].conservedVariables, states[key]), key


def test_stagewise_3d_march_matches_whole_run():
    """A 3-D periodic viscous box on the z marches, stage by stage through the C ABI --
    forward substeps (the march rotates three state buffers: stages 2 and 3 read the state
    at the start of the step where stage 1 left it) and adjoint substeps over migrateTo
    from the step store -- against the whole runs, bit for bit."""
    from magudi_amd.solver import Region, RK4Integrator, Solver, UniformCheckpointer
    from test_gpu_parity import periodic_box, perturbed
    c = periodic_box((32, 16, 16), dissipation=False)
    nsteps, K = 3, 2
    q0 = perturbed(c, 7)
    lamT = np.random.default_rng(8).standard_normal((c.nv, c.npoints))
    R = Region(c)
    R.states[0].conservedVariables = q0
    S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=K)
    S.runForwardCheckpointed()
    qT_run = R.states[0].conservedVariables
    S.runAdjointCheckpointed(terminal=lamT, want_gradient=False)
    lam0_run = R.states[0].adjointVariables
    ti = RK4Integrator()
    R.states[0].conservedVariables = q0
    time, dt = 0.0, c.dt
    for step in range(1, nsteps + 1):
        for i in range(1, 5):
            if i in (2, 4):
                time = time + dt / 2.0
            ti.substepForward(R, time, dt, step, i)
    assert np.array_equal(R.states[0].conservedVariables, qT_run)
    # adjoint stage by stage from lamT over the stored forward states
    R.states[0].conservedVariables = q0
    S.runForwardCheckpointed()
    assert R.checkpoint_plan["mode"] == "step store"
    R.states[0].adjointVariables = lamT
    R.adjointInitialCondition(keep=True)
    mig = UniformCheckpointer()
    for step in range(nsteps - 1, -1, -1):
        for i in range(4, 0, -1):
            if i == 1:
                mig.migrateTo(R, step, 4)
            else:
                mig.migrateTo(R, step + 1, i - 1)
            R.updateAdjointForcing(no_adjoint_forcing=(step == 0 and i == 1))
            ti.substepAdjoint(R, 0.0, dt, step, i)
    assert np.array_equal(R.states[0].adjointVariables, lam0_run)
