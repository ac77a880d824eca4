"""Checkpoint storage of the generic-kernel path.  Window mode (MGD_OPT_STEP_STORE 0):
snapshots every save_interval steps, and the window cache (MGD_OPT_WINDOW_CACHE) -- the
forward pass keeps the stage states of its last windows and the adjoint pass reads them
instead of recomputing those windows.  Step mode (MGD_OPT_STEP_STORE 1, and the default
where it fits): the forward pass keeps the state at the start of every step and the inner
stage states of its last steps; the adjoint pass recomputes three stages per step not held.
On the 3-D marches the stages write their outputs straight into the store, the step
cache and the recompute buffer.  Every storage plan must give the same J, sensitivity,
gradient and adjoint state bit for bit -- on a 3-D box (z-march kernels, a ragged last window) and on a 2-D monopole forced
onto the generic kernels (actuator and cost target).  The per-launch timer of the z-march
kernel counts the forward and recompute launches and the stage copies."""
import numpy as np
import pytest

from magudi_amd import cases

pytestmark = pytest.mark.gpu


def test_checkpoint_storage_3d_bit_for_bit():
    from magudi_amd.solver import Region, Solver
    from test_gpu_parity import periodic_box, perturbed
    c = periodic_box((32, 16, 16), dissipation=False)
    nsteps, K = 7, 3                        # windows start at 0, 3, 6 (the last one step)
    q0 = perturbed(c, 41)
    lamT = np.random.default_rng(42).standard_normal((c.nv, c.npoints))
    plans = [(0, 0), (0, 1), (0, -1), (1, 0), (1, 1), (1, -1), (-1, -1)]
    out, plans_seen = {}, {}
    for step, cache in plans:
        print(f"storage step={step} cache={cache}", flush=True)
        R = Region(c)
        R.set_step_store(step)
        R.set_window_cache(cache)
        R.states[0].conservedVariables = q0
        S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=K)
        R.kernel_timer(True)
        S.runForwardCheckpointed()
        n_fwd, c_fwd, ms_fwd = R.kernel_timer_read()
        qT = R.states[0].conservedVariables
        S.runAdjointCheckpointed(terminal=lamT, want_gradient=False)
        n_all, c_all, ms_all = R.kernel_timer_read()
        R.kernel_timer(False)
        out[(step, cache)] = (qT, R.states[0].adjointVariables, n_fwd, c_fwd, n_all, c_all)
        plan = R.checkpoint_plan
        assert plan["mode"] == ("windows" if step == 0 else "step store"), plan
        plans_seen[(step, cache)] = plan
        assert ms_fwd > 0 and ms_all >= ms_fwd * (1 - 1e-5)
    ref = out[(0, 0)]
    for key in plans[1:]:
        assert np.array_equal(out[key][0], ref[0]), key
        assert np.array_equal(out[key][1], ref[1]), key
    for key in plans:
        assert out[key][2] == 4 * nsteps, key
    # window mode -- forward: copies into the cached windows (all but the last stage of a
    # window); adjoint: the uncached windows recomputed
    assert out[(0, 0)][3] == 0
    assert plans_seen[(0, -1)] == {"mode": "windows", "held": 3, "first": 0}
    assert plans_seen[(0, 1)] == {"mode": "windows", "held": 1, "first": 6}
    assert out[(0, -1)][3] == 4 * 3 - 1 + 4 * 3 - 1 + 4 * 1
    assert out[(0, -1)][4] == out[(0, -1)][2]           # nothing recomputed
    assert out[(0, 0)][4] == 4 * nsteps + 4 * 3 + 4 * 3 + 4 * 1 - 2
    assert out[(0, 1)][4] == 4 * nsteps + 4 * 3 + 4 * 3 - 2   # the last window was kept
    # step mode -- the marches write the state at the start of each step and the inner
    # stage states of the held steps straight into the store and the step cache (no
    # copies); adjoint: three stages per step not held, into the recompute buffer
    for key, held in (((1, 0), 0), ((1, 1), K), ((1, -1), nsteps), ((-1, -1), nsteps)):
        assert plans_seen[key] == {"mode": "step store", "held": held, "first": nsteps - held}
        assert out[key][3] == 0 and out[key][5] == 0, key
        assert out[key][4] == 4 * nsteps + 3 * (nsteps - held), key


@pytest.mark.parametrize("step,cache", [(0, 1), (0, -1), (1, 0), (1, 2), (-1, -1)])
def test_checkpoint_storage_2d_generic_bit_for_bit(step, cache):
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=41, nsteps=10, save_interval=4)
    c.dt = 0.2
    q0 = c.initial_state()
    res = []
    for st, opt in ((0, 0), (step, cache)):
        R = Region(c, force_plain=True)
        R.set_step_store(st)
        R.set_window_cache(opt)
        R.states[0].conservedVariables = q0
        rng = np.random.default_rng(9)
        ctl = 1e-2 * rng.standard_normal((1, 4 * c.nsteps, R.n_actuator))
        sens, grad, J = Solver(R).runAdjoint(control=ctl)
        res.append((sens, grad, J, R.states[0].adjointVariables))
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_march_shape_changes_no_value():
    """The 3-D marches' work-unit shape -- z chunk 32, 64 or 128 planes, tiles 16 x 16 or
    32 x 16 (threads without a halo point), and the autotuner's choice -- on a 64 x 32 x
    80 box (partial chunks for 32 and 64): forward and adjoint runs bit for bit the same."""
    from magudi_amd.solver import Region, Solver
    from test_gpu_parity import periodic_box, perturbed
    c = periodic_box((64, 32, 80), dissipation=False)
    nsteps, K = 3, 2
    q0 = perturbed(c, 51)
    lamT = np.random.default_rng(52).standard_normal((c.nv, c.npoints))
    out = []
    for mt, zc in ((16, 32), (16, 64), (16, 128), (32, 32), (32, 64), (32, 128), (0, 0)):
        R = Region(c)
        R.set_march_tile(mt)
        R.set_z_chunk(zc)
        R.states[0].conservedVariables = q0
        S = Solver(R, dt=c.dt, nsteps=nsteps, save_interval=K)
        S.runForwardCheckpointed()
        qT = R.states[0].conservedVariables
        S.runAdjointCheckpointed(terminal=lamT, want_gradient=False)
        out.append((qT, R.states[0].adjointVariables))
        fwd, adj = R.march_shape
        assert fwd[0] in (16, 32) and fwd[1] in (32, 64, 128)
        assert adj[0] in (16, 32) and adj[1] in (32, 64, 128)
        if zc:
            assert fwd == (mt, zc) and adj == (mt, zc)
    for qT, lam in out[1:]:
        assert np.array_equal(qT, out[0][0])
        assert np.array_equal(lam, out[0][1])


def test_step_store_kolmogorov_bit_for_bit():
    """The Kolmogorov flow (constant transport coefficients) on 32^3, 6 steps from the
    terminal kinetic energy: the step store recomputing every step -- whose adjoint stages
    1-3 read the stress slots of the recompute instead of running the stress pass --, and
    holding the inner states of the last 2 steps, against snapshots and windows, bit for
    bit."""
    from magudi_amd.solver import Region, Solver
    c = cases.kolmogorov3d(n=32, nsteps=6)
    out = {}
    for step, cache in ((0, 0), (1, 0), (1, 1)):
        print(f"kolmogorov storage step={step} cache={cache}", flush=True)
        R = Region(c)
        R.set_step_store(step)
        R.set_window_cache(cache)
        R.states[0].initialCondition = c.initial_state()
        S = Solver(R, dt=c.dt, nsteps=c.nsteps, save_interval=2)
        R.restart()
        S.runForwardCheckpointed()
        J = R.terminalKineticEnergy()
        S.runAdjointCheckpointed(terminal=True, want_gradient=False)
        out[(step, cache)] = (J, R.states[0].conservedVariables, R.states[0].adjointVariables)
        assert R.checkpoint_plan["held"] == (2 if (step, cache) == (1, 1) else 0)
    for key in ((1, 0), (1, 1)):
        for a, b in zip(out[key], out[(0, 0)]):
            assert np.array_equal(a, b), key
