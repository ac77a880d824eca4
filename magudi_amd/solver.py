"""Host-side mirror of magudi's Region / State / Patch / Solver API over the HIP path.

The names and argument meanings follow the reference types:

* ``Region``   -- t_Region (include/Region.f90): ``setup``, ``computeRhs(mode)``,
                  ``getTimeStepSize``; owns the HBM-resident states of
                  ``n_instances`` independent copies of one grid (batching).
* ``State``    -- t_State (include/State.f90): conserved / adjoint / rhs arrays,
                  accessed with ``region.states[i]`` as host numpy arrays.
* ``Solver``   -- t_Solver (include/Solver.f90): ``runForward``, ``runAdjoint``.

Every call goes through libmagudi_hip.so; errors raise ``MagudiError`` (the
reference's ``gracefulExit``).  Nothing here computes the physics on the CPU.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional

import numpy as np

from . import capi
from .capi import FORWARD, ADJOINT, check, dptr
from .cases import Case

__all__ = ["Region", "State", "Solver", "RK4Integrator", "UniformCheckpointer", "FORWARD",
           "ADJOINT"]


def check_run(region: "Region", rc: int):
    """A run's status: errors raise MagudiError; MGD_SOLUTION_LIMIT marks the region's
    solutionCrashed (the flagged instances' J / sensitivity are HUGE, as in
    checkSolutionLimits, src/SolverImpl.f90:814-819)."""
    region.solutionCrashed = rc == capi.SOLUTION_LIMIT
    if rc != capi.SOLUTION_LIMIT:
        check(rc)


class State:
    """View of one instance's arrays (t_State)."""

    def __init__(self, region: "Region", index: int):
        # a weak reference: the region's device memory goes as soon as the last reference
        # to the region does, not at the next garbage collection
        self._r = weakref.proxy(region)
        self.index = index

    def _get(self, which):
        out = np.empty((self._r.nv, self._r.npoints))
        check(capi.lib().mgd_state_get(self._r.handle, which, self.index, dptr(out)))
        return out

    def _put(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self._r.nv, self._r.npoints)
        check(capi.lib().mgd_state_put(self._r.handle, which, self.index, dptr(a)))

    @property
    def conservedVariables(self):
        return self._get(capi.CONSERVED)

    @conservedVariables.setter
    def conservedVariables(self, v):
        self._put(capi.CONSERVED, v)

    @property
    def adjointVariables(self):
        return self._get(capi.ADJOINT_VARS)

    @adjointVariables.setter
    def adjointVariables(self, v):
        self._put(capi.ADJOINT_VARS, v)

    @property
    def rightHandSide(self):
        return self._get(capi.RHS)

    @property
    def initialCondition(self):
        return self._get(capi.INITIAL_STATE)

    @initialCondition.setter
    def initialCondition(self, v):
        self._put(capi.INITIAL_STATE, v)


class Region:
    """t_Region over the HIP path.  ``Region(case, n_instances)`` performs
    Region%setup, Grid%update and the patch / functional / controller set-up of
    Solver%setup (sponge strengths, far-field penalties, mean pressure)."""

    def __init__(self, case: Case, n_instances: int = 1, force_plain: bool = False,
                 tile_shape: int = -1, resident_states: Optional[bool] = None):
        self.case = case
        self.nd = case.nd
        self.nv = case.nv
        self.npoints = case.npoints
        self.n_instances = n_instances
        d = capi.RegionDesc()
        d.nd = case.nd
        for i in range(3):
            d.n[i] = case.n[i]
            d.periodicity[i] = case.periodicity[i]
            d.periodic_length[i] = case.periodic_length[i]
        d.scheme = case.scheme
        d.viscous = int(case.viscous)
        d.dissipation = int(case.dissipation)
        d.composite_dissipation = int(case.composite_dissipation)
        d.gamma = case.gamma
        d.reynolds = case.reynolds
        d.prandtl = case.prandtl
        d.power_law = case.power_law
        d.bulk_ratio = case.bulk_ratio
        d.dissipation_amount = case.dissipation_amount
        d.n_instances = n_instances
        coords = np.ascontiguousarray(case.coords, dtype=np.float64)
        h = C.c_void_p()
        check(capi.lib().mgd_region_create(C.byref(d), dptr(coords), C.byref(h)))
        self.handle = h
        if force_plain:
            check(capi.lib().mgd_region_set_option(self.handle, capi.OPT_FORCE_PLAIN, 1))
        if tile_shape != -1:
            check(capi.lib().mgd_region_set_option(self.handle, capi.OPT_TILE_SHAPE, tile_shape))
        if resident_states is not None:
            # True: every RK stage state of a checkpointed forward run stays in HBM (no window
            # recompute); False: snapshots and window recompute; None (default): resident
            # where the states fit in the device memory left free
            check(capi.lib().mgd_region_set_option(self.handle, capi.OPT_RESIDENT_STATES,
                                                   1 if resident_states else 0))
        for p in case.patches:
            pd = capi.PatchDesc()
            pd.type = p.type
            pd.normal_dir = p.normal_dir
            for i in range(6):
                pd.extent[i] = p.extent[i]
            for i, v in enumerate(list(p.params)[:4]):
                pd.p[i] = v
            check(capi.lib().mgd_region_add_patch(self.handle, C.byref(pd)))
        for which, arr in ((capi.FIELD_TARGET_STATE, case.target),
                           (capi.FIELD_CONTROL_MOLLIFIER, case.control_mollifier),
                           (capi.FIELD_TARGET_MOLLIFIER, case.target_mollifier)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64)
                check(capi.lib().mgd_region_set_field(self.handle, which, dptr(a)))
        for s in case.sources:
            loc = np.array((list(s.location) + [0.0, 0.0, 0.0])[:3], dtype=np.float64)
            check(capi.lib().mgd_region_add_source(self.handle, dptr(loc), s.amplitude,
                                                   s.frequency, s.radius, s.phase))
        check(capi.lib().mgd_region_finalize(self.handle))
        self.states = [State(self, i) for i in range(n_instances)]
        self.solutionCrashed = False
        act = case.patches_of(3)
        self.n_actuator = len(case.patch_points(act[0])) if act else 0

    def close(self):
        if getattr(self, "handle", None):
            capi.lib().mgd_region_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- t_Region procedures --------------------------------------------------
    def computeRhs(self, mode: int, time: float = 0.0, control: Optional[np.ndarray] = None,
                   forcing_factor: float = 1.0):
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        check(capi.lib().mgd_compute_rhs(self.handle, mode, time, dptr(ctl), forcing_factor))

    def restart(self):
        """Conserved state of every instance := its resident initial condition (device copy)."""
        check(capi.lib().mgd_state_copy(self.handle, capi.INITIAL_STATE, capi.CONSERVED))

    def sync(self):
        check(capi.lib().mgd_sync(self.handle))

    def updateAdjointForcing(self, no_adjoint_forcing: bool = False):
        check(capi.lib().mgd_update_adjoint_forcing(self.handle, int(no_adjoint_forcing)))

    def getTimeStepSize(self, cfl: Optional[float] = None) -> float:
        """The case's fixed time_step_size, or with ``cfl`` (use_constant_CFL_mode) the
        smallest computeTimeStepSize over the instances' current states."""
        if cfl is None:
            return self.case.dt
        return float(self.computeTimeStepSize(cfl).min())

    def computeTimeStepSize(self, cfl: float) -> np.ndarray:
        """State%computeTimeStepSize of each instance's conserved state (on the device)."""
        out = np.zeros(self.n_instances)
        check(capi.lib().mgd_time_step_size(self.handle, cfl, dptr(out)))
        return out

    def computeCfl(self, dt: float) -> np.ndarray:
        """State%computeCfl of each instance's conserved state at time step dt."""
        out = np.zeros(self.n_instances)
        check(capi.lib().mgd_cfl(self.handle, dt, dptr(out)))
        return out

    def functional(self) -> np.ndarray:
        out = np.zeros(self.n_instances)
        check(capi.lib().mgd_functional(self.handle, dptr(out)))
        return out

    def setSolutionLimits(self, density_range, temperature_range, enable: bool = True):
        """enable_solution_limits with (minimum, maximum) density and temperature: every
        forward stage of a run then checks its input state (checkSolutionLimits)."""
        rng = np.array(list(density_range) + list(temperature_range), dtype=np.float64)
        check(capi.lib().mgd_region_set_solution_limits(self.handle, int(enable), dptr(rng)))

    def checkSolutionLimits(self) -> np.ndarray:
        """Per instance: whether the current conserved state reaches a solution limit."""
        out = (C.c_int * self.n_instances)()
        rc = capi.lib().mgd_check_solution_limits(self.handle, out)
        if rc != capi.SOLUTION_LIMIT:
            check(rc)
        return np.array(list(out), dtype=bool)

    def terminalKineticEnergy(self) -> np.ndarray:
        """Kinetic energy J = sum_p H_p |rho u|^2 / (2 rho) of each instance's current
        conserved state, computed on the device, which also writes dJ/dq as each
        instance's adjoint state (the terminal condition for
        ``Solver.runAdjointCheckpointed(terminal=True)``)."""
        out = np.zeros(self.n_instances)
        check(capi.lib().mgd_terminal_cost(self.handle, capi.TERMINAL_KINETIC_ENERGY, dptr(out)))
        return out

    def controlGradient(self):
        """(gradient, sensitivity) at the current adjoint state: per instance and actuator
        point lambda_E times the control mollifier (ThermalActuator%updateGradient), and per
        instance its H-norm square (computeSensitivity)."""
        grad = np.zeros((self.n_instances, max(self.n_actuator, 1)))
        sens = np.zeros(self.n_instances)
        check(capi.lib().mgd_control_gradient(self.handle, dptr(grad), dptr(sens)))
        return grad[:, :self.n_actuator], sens

    def adjointInitialCondition(self, keep: bool = False):
        """loadInitialCondition(ADJOINT) after a checkpointed forward run: zero (or the
        current adjoint state, keep=True) plus the cost-target forcing of the final state."""
        check(capi.lib().mgd_adjoint_initial_condition(self.handle, int(keep)))

    def grid(self, which: int) -> np.ndarray:
        size = {0: self.nd * self.nd, 1: 1, 2: 1, 3: self.nd}[which]
        out = np.zeros((size, self.npoints))
        check(capi.lib().mgd_region_get_grid(self.handle, which, dptr(out)))
        return out

    def applyOperator(self, op: int, direction: int, x: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.float64).copy()
        ncomp = a.size // self.npoints
        check(capi.lib().mgd_stencil_apply(self.handle, op, direction, dptr(a), ncomp))
        return a

    @property
    def tile_shape(self) -> int:
        """Tile shape the fused 2-D kernels use (-1: generic kernels)."""
        v = C.c_int(0)
        check(capi.lib().mgd_region_get_tile_shape(self.handle, C.byref(v)))
        return v.value

    def time_kernel(self, which: int, reps: int = 20):
        """(ms per launch, algorithmic bytes per launch) of one kernel, HIP-event timed."""
        ms = C.c_double(0.0)
        b = C.c_double(0.0)
        check(capi.lib().mgd_time_kernel(self.handle, which, reps, C.byref(ms), C.byref(b)))
        return ms.value, b.value

    def set_window_cache(self, n: int):
        """MGD_OPT_WINDOW_CACHE: -1 keeps as many windows of the checkpointed forward pass
        as device memory allows (default), 0 none, n > 0 at most n."""
        check(capi.lib().mgd_region_set_option(self.handle, capi.OPT_WINDOW_CACHE, int(n)))

    def set_step_store(self, mode: int):
        """MGD_OPT_STEP_STORE: 1 keeps the state at the start of every step of the
        checkpointed forward pass (and the inner stage states of its last steps, as memory
        and set_window_cache allow), so that the adjoint pass recomputes three stages per
        step not held instead of whole windows; 0 snapshots and windows; -1 (default) step
        store where it fits and the window cache is at its default."""
        check(capi.lib().mgd_region_set_option(self.handle, capi.OPT_STEP_STORE, int(mode)))

    def set_z_chunk(self, zc: int):
        """MGD_OPT_Z_CHUNK: 0 autotunes the 3-D marches' z chunk at the first run, or 32,
        64, 128 fixes it."""
        check(capi.lib().mgd_region_set_option(self.handle, capi.OPT_Z_CHUNK, int(zc)))

    def set_march_tile(self, width: int):
        """MGD_OPT_MARCH_TILE: 0 autotunes the 3-D marches' tile width with the z chunk, or
        16 / 32 fixes it."""
        check(capi.lib().mgd_region_set_option(self.handle, capi.OPT_MARCH_TILE, int(width)))

    @property
    def march_shape(self):
        """((tile width, z chunk) of the forward march, (tile width, z chunk) of the adjoint
        marches) in use (after the autotuner, if it ran)."""
        out = (C.c_int * 4)()
        check(capi.lib().mgd_region_get_march_shape(self.handle, out))
        return (out[0], out[1]), (out[2], out[3])

    @property
    def checkpoint_plan(self) -> dict:
        """Storage plan of the last checkpointed forward run: {"mode": "step store",
        "held": steps whose inner stage states are kept, "first": the first of them},
        {"mode": "windows", "held": windows kept, "first": their first step} (the generic
        kernels) or {"mode": "fused", "held": resident stage states (0: window recompute)}
        (the fused 2-D path without the step store)."""
        out = (C.c_int * 3)()
        check(capi.lib().mgd_region_get_checkpoint_plan(self.handle, out))
        mode = {0: "windows", 1: "step store", 2: "fused"}[out[0]]
        return {"mode": mode, "held": out[1], "first": out[2]}

    @property
    def march_tile(self) -> int:
        """The forward march's tile width in use."""
        return self.march_shape[0][0]

    @property
    def z_chunk(self) -> int:
        """The forward march's z chunk in use."""
        return self.march_shape[0][1]

    def kernel_timer(self, on: bool, which: int = None):
        """Starts (on) or stops per-launch HIP-event timing of the z-march forward kernel
        inside the runs."""
        which = capi.KERNEL_FWD_MARCH3D if which is None else which
        check(capi.lib().mgd_kernel_timer(self.handle, which, 1 if on else 0))

    def kernel_timer_read(self):
        """(launches, launches that also wrote a window copy, summed ms) since the timer
        started."""
        n, nc, ms = C.c_long(0), C.c_long(0), C.c_double(0.0)
        check(capi.lib().mgd_kernel_timer_read(self.handle, C.byref(n), C.byref(nc), C.byref(ms)))
        return n.value, nc.value, ms.value

    def timings(self):
        out = np.zeros(3)
        n = (C.c_int * 3)()
        check(capi.lib().mgd_last_timings(self.handle, dptr(out), n))
        return out, list(n)


class RK4Integrator:
    """t_RK4Integrator (include/RK4Integrator.f90, src/RK4IntegratorImpl.f90:65-270): one RK4
    stage at a time on the region's HBM-resident state; the RK buffers live in the region.
    ``norm`` holds the quadrature weights b_i of the stages."""

    nStages = 4
    norm = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)

    def substepForward(self, region: "Region", time: float, timeStepSize: float, timestep: int,
                       stage: int, control: Optional[np.ndarray] = None):
        """Stage ``stage`` (1..4) of the forward step; ``time`` is State%time of the stage
        (t, t + dt/2, t + dt/2, t + dt), ``control`` the stage's actuator forcing per
        instance and actuator point."""
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        check(capi.lib().mgd_rk4_substep(region.handle, FORWARD, stage, time, timeStepSize,
                                         dptr(ctl)))

    def substepAdjoint(self, region: "Region", time: float, timeStepSize: float, timestep: int,
                       stage: int):
        """Stage ``stage`` (4..1) of the reversed step, about the forward stage state in the
        conserved variables (UniformCheckpointer.migrateTo puts it there)."""
        check(capi.lib().mgd_rk4_substep(region.handle, ADJOINT, stage, time, timeStepSize, None))


class UniformCheckpointer:
    """t_UniformCheckpointer's migrateTo (src/UniformCheckpointerImpl.f90:78-208) over the HBM
    checkpoints of the region's last Solver.runForwardCheckpointed."""

    def migrateTo(self, region: "Region", timestep: int, stage: int):
        check(capi.lib().mgd_checkpoint_migrate(region.handle, timestep, stage))


class Solver:
    """t_Solver: time marching over a Region (RK4, uniform checkpointing in HBM)."""

    def __init__(self, region: Region, dt: Optional[float] = None, nsteps: Optional[int] = None,
                 save_interval: Optional[int] = None):
        self.region = region
        self.dt = dt if dt is not None else region.case.dt
        self.nTimesteps = nsteps if nsteps is not None else region.case.nsteps
        self.saveInterval = save_interval if save_interval is not None else region.case.save_interval

    def runForward(self, control: Optional[np.ndarray] = None, t0: float = 0.0,
                   use_graph: bool = False) -> np.ndarray:
        """Solver%runForward from the current conserved state; returns J per instance."""
        r = self.region
        J = np.zeros(r.n_instances)
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        check_run(r, capi.lib().mgd_run_forward(r.handle, t0, self.dt, self.nTimesteps, dptr(ctl),
                                         dptr(J), int(use_graph)))
        return J

    def runAdjoint(self, control: Optional[np.ndarray] = None, t0: float = 0.0,
                   want_gradient: bool = True, use_graph: bool = False):
        """Solver%runAdjoint (forward pass with HBM checkpoints, then the discrete adjoint).

        Returns (sensitivity, gradient, J): sensitivity and J per instance; gradient per
        instance, per RK stage in forward order, per actuator patch point."""
        r = self.region
        J = np.zeros(r.n_instances)
        sens = np.zeros(r.n_instances)
        grad = None
        if want_gradient and r.n_actuator > 0:
            grad = np.zeros((r.n_instances, 4 * self.nTimesteps, r.n_actuator))
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        check_run(r, capi.lib().mgd_run_adjoint(r.handle, t0, self.dt, self.nTimesteps, self.saveInterval,
                                         dptr(ctl), dptr(grad), dptr(sens), dptr(J),
                                         int(use_graph)))
        return sens, grad, J

    def runForwardCheckpointed(self, control: Optional[np.ndarray] = None,
                               t0: float = 0.0) -> np.ndarray:
        """First half of Solver%runAdjoint (src/SolverImpl.f90:914-1060): the forward run
        from the current conserved state with HBM checkpoints (Region.checkpoint_plan).
        Returns J (the time-integrated functional) per instance; afterwards the
        conserved state is the final state."""
        r = self.region
        J = np.zeros(r.n_instances)
        ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
        check_run(r, capi.lib().mgd_run_forward_checkpointed(r.handle, t0, self.dt, self.nTimesteps,
                                                      self.saveInterval, dptr(ctl), dptr(J)))
        return J

    def runAdjointCheckpointed(self, terminal: Optional[np.ndarray] = None,
                               want_gradient: bool = True):
        """Second half of Solver%runAdjoint: the discrete adjoint marched back from the
        checkpoints of the last runForwardCheckpointed.  ``terminal`` is the adjoint state at
        the final time per instance, shape (n_instances, nv, N) or (nv, N) for one
        instance -- the reference's adjoint_nonzero_initial_condition
        (src/SolverImpl.f90:368-428, 1096-1102), e.g. the gradient of a cost on the final
        state --, True for the adjoint state already on the device (as
        ``Region.terminalKineticEnergy`` leaves it), or None for a zero terminal state.  Returns (sensitivity, gradient) as
        runAdjoint does; afterwards each instance's adjoint state is the adjoint at t0."""
        r = self.region
        if terminal is not None and terminal is not True:
            t = np.asarray(terminal, dtype=np.float64).reshape(r.n_instances, r.nv, r.npoints)
            for i in range(r.n_instances):
                r.states[i].adjointVariables = t[i]
        sens = np.zeros(r.n_instances)
        grad = None
        if want_gradient and r.n_actuator > 0:
            grad = np.zeros((r.n_instances, 4 * self.nTimesteps, r.n_actuator))
        check(capi.lib().mgd_run_adjoint_checkpointed(r.handle, 0 if terminal is None else 1,
                                                      dptr(grad), dptr(sens)))
        return sens, grad
