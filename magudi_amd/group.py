"""Uneven batches: independent regions of any size and problem type run together.

Two ways, with the same results as each region run alone:

* concurrent runs: each region owns its HIP streams, so the runs of several regions
  overlap on the GPU when they are issued from different host threads (the C ABI calls
  release the Python interpreter lock); each region keeps its own tile shape;
* one launch per RK stage for all members (``run_adjoint_merged``,
  ``mgd_group_run_adjoint``): the fused 2-D members of equal step count march in lock
  step, every launch covering the 16 x 16 tiles of all of them; the other members run
  concurrently afterwards.  Many small grids gain most (launch-bound alone).
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import capi
from .cases import Case
from .capi import check, dptr
from .solver import Region, Solver

__all__ = ["RegionGroup"]


class RegionGroup:
    """``shooting`` maps a member index to the time_splitting options of a multiple-shooting
    member (nsplit, nts, penalty_weight, and optionally penalty_type, huber_threshold,
    save_interval, ics, t0s): that member runs msforward + msadjoint over its nsplit time
    segments, the segments being the instances of its region (bin/msforward.f90:242-365,
    bin/msadjoint.f90:287-384); without ics / t0s the segments start from
    multishoot.segment_initial_guesses."""

    def __init__(self, cases: Sequence[Case], n_instances: Optional[Sequence[int]] = None,
                 force_plain: bool = False, workers: int = 8, shooting: Optional[dict] = None):
        from .multishoot import MultipleShooting, segment_initial_guesses
        n_instances = list(n_instances) if n_instances is not None else [1] * len(cases)
        self.cases = list(cases)
        self.shooting = {}
        self.shooting_results = {}      # member index -> msadjoint's dict of its last run
        self.regions, self.solvers = [], []
        for k, (c, ni) in enumerate(zip(self.cases, n_instances)):
            opt = (shooting or {}).get(k)
            if opt is None:
                r = Region(c, n_instances=ni, force_plain=force_plain)
                for i in range(r.n_instances):
                    r.states[i].initialCondition = c.initial_state()
                self.regions.append(r)
                self.solvers.append(Solver(r))
                continue
            opt = dict(opt)
            ics, t0s = opt.pop("ics", None), opt.pop("t0s", None)
            if ics is None:
                ics, t0s = segment_initial_guesses(c, opt["nsplit"], opt["nts"])
            M = MultipleShooting(c, force_plain=force_plain, **opt)
            self.shooting[k] = (M, ics, t0s)
            self.regions.append(M.region)
            self.solvers.append(Solver(M.region, dt=M.dt, nsteps=M.nts, save_interval=M.K))
        self.workers = max(1, workers)

    @property
    def grid_point_steps(self) -> int:
        return sum(r.npoints * r.n_instances * s.nTimesteps for r, s in zip(self.regions, self.solvers))

    def _shooting_run(self, k, control, want_gradient):
        """msforward + msadjoint of a multiple-shooting member: (total sensitivity of the
        objective -- the sum of the segments' control inner products and IC-gradient inner
        products --, control gradient, J of the objective)."""
        M, ics, t0s = self.shooting[k]
        J = M.msforward(ics, t0s, control=control)
        out = M.msadjoint(want_control_gradient=want_gradient)
        self.shooting_results[k] = out
        return (np.array([out["total"]]), out["control_grad"], np.array([J]))

    def _map(self, fn):
        with cf.ThreadPoolExecutor(max_workers=min(self.workers, len(self.regions))) as ex:
            return list(ex.map(fn, range(len(self.regions))))

    def restart(self):
        for k, r in enumerate(self.regions):
            if k not in self.shooting:   # msforward restarts its segments itself
                r.restart()

    def run_adjoint(self, controls: Optional[List[Optional[np.ndarray]]] = None,
                    want_gradient: bool = True):
        """Solver%runAdjoint of every region (msforward + msadjoint of the shooting members),
        concurrently; list of (sens, grad, J)."""
        controls = controls or [None] * len(self.regions)

        def one(k):
            if k in self.shooting:
                return self._shooting_run(k, controls[k], want_gradient)
            return self.solvers[k].runAdjoint(control=controls[k], want_gradient=want_gradient)
        return self._map(one)

    def _merged_members(self) -> List[int]:
        """Members the one-launch path takes: fused 2-D regions (16 x 16 tiles fit) with
        the step count and snapshot interval of the first of them."""
        idx = [k for k, r in enumerate(self.regions)
               if r.nd == 2 and r.tile_shape >= 0 and k not in self.shooting]
        if not idx:
            return []
        s0 = self.solvers[idx[0]]
        return [k for k in idx if self.solvers[k].nTimesteps == s0.nTimesteps
                and self.solvers[k].saveInterval == s0.saveInterval]

    def _group_handle(self, members: List[int]):
        key = tuple(members)
        if getattr(self, "_group_key", None) != key:
            self.close()
            arr = (C.c_void_p * len(members))(*[self.regions[k].handle.value for k in members])
            h = C.c_void_p()
            check(capi.lib().mgd_group_create(arr, len(members), C.byref(h)))
            self._group, self._group_key = h, key
        return self._group

    def close(self):
        if getattr(self, "_group", None) is not None:
            capi.lib().mgd_group_destroy(self._group)
            self._group = None
            self._group_key = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _run_adjoint_merged(self, members, controls, want_gradient):
        """mgd_group_run_adjoint over `members`: list of (sens, grad, J) per member."""
        s0 = self.solvers[members[0]]
        outs, keep = [], []
        n = len(members)
        dts = np.array([self.solvers[k].dt for k in members], dtype=np.float64)
        ctl = (C.POINTER(C.c_double) * n)()
        grd = (C.POINTER(C.c_double) * n)()
        sen = (C.POINTER(C.c_double) * n)()
        jj = (C.POINTER(C.c_double) * n)()
        for j, k in enumerate(members):
            r = self.regions[k]
            J = np.zeros(r.n_instances)
            sens = np.zeros(r.n_instances)
            grad = None
            if want_gradient and r.n_actuator > 0:
                grad = np.zeros((r.n_instances, 4 * s0.nTimesteps, r.n_actuator))
            c = None if controls[k] is None else np.ascontiguousarray(controls[k], dtype=np.float64)
            keep.append(c)
            ctl[j], grd[j], sen[j], jj[j] = dptr(c), dptr(grad), dptr(sens), dptr(J)
            outs.append((sens, grad, J))
        check(capi.lib().mgd_group_run_adjoint(self._group_handle(members), 0.0, dptr(dts),
                                               s0.nTimesteps, s0.saveInterval, ctl,
                                               grd if want_gradient else None, sen, jj))
        return outs

    def run_adjoint_merged(self, controls: Optional[List[Optional[np.ndarray]]] = None,
                           want_gradient: bool = True):
        """Solver%runAdjoint of every region: the fused 2-D members of equal step count in
        one launch per stage (mgd_group_run_adjoint), the others concurrently beside them;
        list of (sens, grad, J) in member order."""
        controls = controls or [None] * len(self.regions)
        members = self._merged_members()
        rest = [k for k in range(len(self.regions)) if k not in members]
        results = [None] * len(self.regions)

        def one(k):
            if k in self.shooting:
                return self._shooting_run(k, controls[k], want_gradient)
            return self.solvers[k].runAdjoint(control=controls[k], want_gradient=want_gradient)
        # the merged launches first, then the other members concurrently: issued at the
        # same time, their many small launches would share the hardware queues with the
        # merged ones and hold them up
        if members:
            for k, o in zip(members, self._run_adjoint_merged(members, controls, want_gradient)):
                results[k] = o
        if rest:
            with cf.ThreadPoolExecutor(max_workers=max(1, min(self.workers, len(rest)))) as ex:
                for k, o in zip(rest, ex.map(one, rest)):
                    results[k] = o
        return results

    def run_forward(self, controls: Optional[List[Optional[np.ndarray]]] = None):
        controls = controls or [None] * len(self.regions)

        def one(k):
            return self.solvers[k].runForward(control=controls[k])
        return self._map(one)

    def sync(self):
        for r in self.regions:
            r.sync()
