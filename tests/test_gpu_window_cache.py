"""The window cache of checkpointed runs on the generic-kernel path (MGD_OPT_WINDOW_CACHE):
the forward pass keeps the stage states of its last windows and the adjoint pass reads
them instead of recomputing those windows.  Runs with no cache, a partial cache and the
default (every window, on these small grids) must give the same J, sensitivity, gradient
and adjoint state bit for bit -- on a 3-D box (z-march kernels, a ragged last window) and
on a 2-D monopole forced onto the generic kernels (actuator and cost target).  The
per-launch timer of the z-march kernel counts the forward and recompute launches."""
import numpy as np
import pytest

from magudi_amd import cases

pytestmark = pytest.mark.gpu


def test_window_cache_3d_bit_for_bit():
    from magudi_amd.solver import Region, Solver
    from test_gpu_parity import periodic_box, perturbed
    c = periodic_box((32, 16, 16), dissipation=False)
    nsteps, K = 7, 3                        # windows start at 0, 3, 6 (the last one step)
    q0 = perturbed(c, 41)
    lamT = np.random.default_rng(42).standard_normal((c.nv, c.npoints))
    out = {}
    for cache in (0, 1, -1):
        R = Region(c)
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
        out[cache] = (qT, R.states[0].adjointVariables, n_fwd, c_fwd, n_all, c_all)
        assert ms_fwd > 0 and ms_all >= ms_fwd * (1 - 1e-5)
    for cache in (1, -1):
        assert np.array_equal(out[cache][0], out[0][0])
        assert np.array_equal(out[cache][1], out[0][1]), cache
    # forward: 4 launches per step; copies into the cached windows (all but the last
    # stage of a window).  Recompute: the uncached windows of the adjoint pass.
    assert out[0][2] == 4 * nsteps and out[0][3] == 0
    assert out[-1][2] == 4 * nsteps and out[-1][3] == 4 * 3 - 1 + 4 * 3 - 1 + 4 * 1
    assert out[-1][4] == out[-1][2]                      # nothing recomputed
    assert out[0][4] == 4 * nsteps + 4 * 3 + 4 * 3 + 4 * 1 - 2
    assert out[1][4] == 4 * nsteps + 4 * 3 + 4 * 3 - 2   # the last window was kept


@pytest.mark.parametrize("cache", [1, -1])
def test_window_cache_2d_generic_bit_for_bit(cache):
    from magudi_amd.solver import Region, Solver
    c = cases.acoustic_monopole(n=41, nsteps=10, save_interval=4)
    c.dt = 0.2
    q0 = c.initial_state()
    res = []
    for opt in (0, cache):
        R = Region(c, force_plain=True)
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
